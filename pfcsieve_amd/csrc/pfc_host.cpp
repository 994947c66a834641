// pfc_host.cpp -- host driver of libpfcsieve.so: devices, jobs, start-residue
// tables, the segment loop and the factor harvest (include/pfc.h).
//
// This is the device-independent half of the reference's cl_sieve.cpp
// (setupSearch 596-665, tables 824-1326, main loop 1328-1761, getResults
// 450-593) and verifyprime.cpp, rebuilt over HIP.  No call exits the process.
#include <hip/hip_runtime.h>

#include <errno.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/pfc.h"
#include "pfc_internal.h"
#include "pfc_share.h"

using pfc::set_error;
using pfc::hip_fail;

typedef unsigned __int128 u128;

// ====================================================================== primes

namespace {

// Primes in [lo, hi] (inclusive, hi < 2^32) by a segmented odd-only sieve of
// Eratosthenes.  Stands in for primesieve_generate_primes (cl_sieve.cpp:850,
// 996, 1000, 1158, 1498): prime lists are unique, so any exact sieve agrees.
std::vector<uint32_t> primes_in(uint64_t lo, uint64_t hi) {
    std::vector<uint32_t> out;
    if (hi < 2 || lo > hi) return out;
    if (lo <= 2) out.push_back(2);
    uint64_t root = (uint64_t)sqrt((double)hi);
    while (root * root > hi) --root;
    while ((root + 1) * (root + 1) <= hi) ++root;
    std::vector<uint8_t> small(root + 1, 0);
    std::vector<uint32_t> base;
    for (uint64_t i = 3; i <= root; i += 2) {
        if (small[i]) continue;
        base.push_back((uint32_t)i);
        for (uint64_t j = i * i; j <= root; j += 2 * i) small[j] = 1;
    }
    const uint64_t SEG = 1u << 21;  // odd numbers per segment
    std::vector<uint8_t> seg(SEG);
    uint64_t start = lo < 3 ? 3 : (lo | 1);
    for (uint64_t s = start; s <= hi; s += 2 * SEG) {
        const uint64_t e = std::min<uint64_t>(hi, s + 2 * SEG - 2);  // last odd
        const uint64_t cnt = (e - s) / 2 + 1;
        std::fill(seg.begin(), seg.begin() + cnt, 0);
        for (uint32_t q : base) {
            const uint64_t q2 = (uint64_t)q * q;
            if (q2 > e) break;
            uint64_t f = q2 >= s ? q2 : ((s + q - 1) / q) * q;
            if (!(f & 1)) f += q;
            for (uint64_t x = f; x <= e; x += 2 * q) seg[(x - s) / 2] = 1;
        }
        for (uint64_t k = 0; k < cnt; ++k)
            if (!seg[k]) out.push_back((uint32_t)(s + 2 * k));
    }
    return out;
}

// Setup table (include/pfc.h pfc_setup_table): exponents of every prime r <= N
// in X(N) = N! (kind 0), N# (kind 1) or N!/N# (kind 2), written in base
// 2^digit_bits; group (b, t) holds the primes with digit t at position b,
// multiplied into chunks below 2^63 (so that a chunk product accepts an
// unreduced accumulator).
struct HostTable {
    std::vector<uint64_t> chunks;
    std::vector<uint32_t> groups;  // npos * (2^digit_bits - 1) + 1 offsets
    uint32_t npos = 0;             // digit positions of the largest exponent
    uint32_t digit_bits = 1;
    uint32_t kind = 0, N = 0;      // X(N) = N!, N# or N!/N#
};

constexpr u128 kChunkLimit = (u128)1 << 63;
// Widest n-range (nmax - nmin) whose composites are also listed for the walk
// (4 bytes each); wider C and F|C ranges walk the bitmap alone.
constexpr uint32_t kMaxCompositeList = 1u << 27;
constexpr uint32_t kMaxDigitBits = 8;

uint32_t positions_of(uint32_t maxe, uint32_t k) {
    uint32_t n = 0;
    while (n * k < 32 && (maxe >> (n * k))) ++n;
    return n;
}

// Modular products the set-up spends on a table of digit width k, estimated
// from the bits of each group (about 60 of a chunk's 63 are filled), plus the
// suffix products (the largest non-empty digit of each position, k > 1) and
// the squarings between positions.
double table_cost(const std::vector<uint32_t>& pr, const std::vector<uint32_t>& ex,
                  const std::vector<float>& lg, uint32_t maxe, uint32_t k) {
    const uint32_t npos = positions_of(maxe, k), B = 1u << k;
    std::vector<double> bits((size_t)npos * B, 0.0);
    std::vector<uint32_t> top(npos, 0);
    for (size_t i = 0; i < pr.size(); ++i) {
        for (uint32_t b = 0; b < npos; ++b) {
            const uint32_t t = (ex[i] >> (b * k)) & (B - 1);
            if (!t) {
                if ((ex[i] >> (b * k)) == 0) break;
                continue;
            }
            bits[(size_t)b * B + t] += lg[i];
            top[b] = std::max(top[b], t);
        }
    }
    double cost = (double)k * (npos ? npos - 1 : 0);
    for (double x : bits) cost += ceil(x / 60.0);
    if (k > 1)
        for (uint32_t t : top) cost += t;
    return cost;
}

HostTable build_table(uint32_t kind, uint32_t N) {
    HostTable T;
    T.kind = kind;
    T.N = N;
    std::vector<uint32_t> pr = primes_in(2, N);
    std::vector<uint32_t> ex(pr.size());
    std::vector<float> lg(pr.size());
    uint32_t maxe = 0;
    for (size_t i = 0; i < pr.size(); ++i) {
        uint32_t e = 1;
        if (kind != 1) {  // Legendre: sum floor(N / r^k)  (cl_sieve.cpp:824-839)
            e = 0;
            for (uint64_t pk = pr[i]; pk <= N; pk *= pr[i]) e += (uint32_t)(N / pk);
            if (kind == 2) e -= 1;
        }
        ex[i] = e;
        lg[i] = log2f((float)pr[i]);
        maxe = std::max(maxe, e);
    }
    // Digit width: the cheapest by the estimate (ties to the smaller).
    uint32_t k = 1;
    if (maxe > 1) {
        double best = table_cost(pr, ex, lg, maxe, 1);
        for (uint32_t kk = 2; kk <= kMaxDigitBits; ++kk) {
            const double c = table_cost(pr, ex, lg, maxe, kk);
            if (c < best) {
                best = c;
                k = kk;
            }
        }
    }
    T.digit_bits = k;
    T.npos = positions_of(maxe, k);
    const uint32_t B = 1u << k, ndig = B - 1;
    // Buckets in group order (position-major from the highest position,
    // digits from B-1 down), each filled in increasing prime order.
    std::vector<std::vector<uint32_t>> bucket((size_t)T.npos * ndig);
    for (size_t i = 0; i < pr.size(); ++i) {
        for (uint32_t b = 0; b < T.npos; ++b) {
            const uint32_t rest = ex[i] >> (b * k);
            if (!rest) break;
            const uint32_t t = rest & (B - 1);
            if (t) bucket[(size_t)(T.npos - 1 - b) * ndig + (B - 1 - t)].push_back(pr[i]);
        }
    }
    T.groups.assign(bucket.size() + 1, 0);
    // Chunks are filled largest-first: the largest remaining primes while they
    // fit, then the smallest remaining ones as filler (about 4% fewer chunks
    // than filling in increasing order; the group's product is unchanged).
    for (size_t g = 0; g < bucket.size(); ++g) {
        T.groups[g] = (uint32_t)T.chunks.size();
        const std::vector<uint32_t>& v = bucket[g];
        size_t lo = 0, hi = v.size();
        while (lo < hi) {
            uint64_t acc = v[--hi];
            while (lo < hi && (u128)acc * v[hi - 1] < kChunkLimit) acc *= v[--hi];
            while (lo < hi && (u128)acc * v[lo] < kChunkLimit) acc *= v[lo++];
            T.chunks.push_back(acc);
        }
    }
    T.groups[bucket.size()] = (uint32_t)T.chunks.size();
    return T;
}

// 64-bit Montgomery helpers for the host-side isPrime (R = 2^64).
inline uint64_t h_inv(uint64_t p) {
    uint64_t x = p;  // p*p == 1 mod 8 gives 3 bits; each step doubles them
    for (int i = 0; i < 5; ++i) x *= 2 - p * x;
    return x;
}
inline uint64_t h_redc(u128 t, uint64_t p, uint64_t pinv) {
    const uint64_t m = (uint64_t)t * pinv;
    const uint64_t mp = (uint64_t)(((u128)m * p) >> 64);
    const uint64_t hi = (uint64_t)(t >> 64);
    return hi >= mp ? hi - mp : hi - mp + p;
}

}  // namespace

// ====================================================================== device

// Device memory, streams and small pinned blocks kept for reuse by later
// jobs.  Creating and destroying a job otherwise costs about 4 ms of runtime
// calls (measured on MI355X: hipStreamCreate 1.1 ms, hipStreamDestroy 0.5 ms,
// hipFree 0.1-0.3 ms per buffer, hipHostFree 0.25 ms), which batched and
// served job streams pay for every job.  Device buffers are cached by size
// class (at most 1/8 above the request), up to kPoolCacheBytes; a failed
// allocation empties the cache and is tried once more.
struct DevicePool {
    std::mutex mu;
    std::multimap<size_t, void*> buffers;   // size class -> free buffer
    size_t cached = 0;                      // bytes in `buffers`
    std::vector<hipStream_t> streams;       // idle, synchronised streams
    std::vector<void*> pinned;              // free kPinnedBlock-byte pinned blocks
    std::vector<void*> slabs;               // pinned slabs the blocks come from
};
constexpr size_t kPoolCacheBytes = 16ull << 30;
constexpr size_t kPinnedBlock = 256, kPinnedSlab = 64u << 10;

struct pfc_device {
    int ordinal = 0;
    hipDeviceProp_t prop{};
    pfc_tuning tuning{};
    mutable std::mutex mu;   // guards tuning
    DevicePool pool;
};

namespace {

size_t size_class(size_t bytes) {
    if (bytes <= 4096) return 4096;
    const int b = 63 - __builtin_clzll(bytes);          // 2^b <= bytes < 2^(b+1)
    const size_t step = (size_t)1 << (b - 3);
    return (bytes + step - 1) / step * step;
}

void pool_flush(pfc_device* dev) {
    std::multimap<size_t, void*> b;
    {
        std::lock_guard<std::mutex> g(dev->pool.mu);
        b.swap(dev->pool.buffers);
        dev->pool.cached = 0;
    }
    for (auto& x : b) (void)hipFree(x.second);
}

// A device buffer of at least `bytes`; *cls receives its size class, which
// pool_free needs back.
int pool_malloc(pfc_device* dev, void** p, size_t bytes, size_t* cls) {
    const size_t c = size_class(bytes);
    *cls = c;
    {
        std::lock_guard<std::mutex> g(dev->pool.mu);
        auto it = dev->pool.buffers.find(c);
        if (it != dev->pool.buffers.end()) {
            *p = it->second;
            dev->pool.buffers.erase(it);
            dev->pool.cached -= c;
            return PFC_OK;
        }
    }
    hipError_t e = hipMalloc(p, c);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        pool_flush(dev);
        e = hipMalloc(p, c);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return set_error(PFC_ERR_NOMEM, "device allocation failed");
    }
    return PFC_OK;
}

// Returns a buffer that no queued work uses any more.
void pool_free(pfc_device* dev, void* p, size_t cls) {
    if (!p) return;
    {
        std::lock_guard<std::mutex> g(dev->pool.mu);
        if (dev->pool.cached + cls <= kPoolCacheBytes) {
            dev->pool.buffers.emplace(cls, p);
            dev->pool.cached += cls;
            return;
        }
    }
    (void)hipFree(p);
}

int pool_stream(pfc_device* dev, hipStream_t* s) {
    {
        std::lock_guard<std::mutex> g(dev->pool.mu);
        if (!dev->pool.streams.empty()) {
            *s = dev->pool.streams.back();
            dev->pool.streams.pop_back();
            return PFC_OK;
        }
    }
    hipError_t e = hipStreamCreateWithFlags(s, hipStreamNonBlocking);
    return e == hipSuccess ? PFC_OK : hip_fail(e, "stream create");
}

void pool_stream_put(pfc_device* dev, hipStream_t s) {
    std::lock_guard<std::mutex> g(dev->pool.mu);
    dev->pool.streams.push_back(s);
}

// A pinned host block of kPinnedBlock bytes.
int pool_pinned(pfc_device* dev, void** p) {
    std::lock_guard<std::mutex> g(dev->pool.mu);
    if (dev->pool.pinned.empty()) {
        void* slab = nullptr;
        hipError_t e = hipHostMalloc(&slab, kPinnedSlab, hipHostMallocDefault);
        if (e != hipSuccess) return set_error(PFC_ERR_NOMEM, "pinned host allocation failed");
        dev->pool.slabs.push_back(slab);
        for (size_t off = 0; off < kPinnedSlab; off += kPinnedBlock)
            dev->pool.pinned.push_back((char*)slab + off);
    }
    *p = dev->pool.pinned.back();
    dev->pool.pinned.pop_back();
    return PFC_OK;
}

void pool_pinned_put(pfc_device* dev, void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> g(dev->pool.mu);
    dev->pool.pinned.push_back(p);
}

}  // namespace

static pfc_tuning tuning_of(const pfc_device* dev) {
    std::lock_guard<std::mutex> g(dev->mu);
    return dev->tuning;
}

static pfc_tuning default_tuning() {
    pfc_tuning t;
    t.primes_per_lane = 1;   // fastest walk shape measured on MI355X (pfc_autotune)
    t.block_size = 256;
    // 2^24-candidate segments give each walk launch 32 rounds of wavefronts
    // per wave slot, so the ramp-up and drain at the ends of a launch cost
    // ~1.5% (2^22: ~6%); 2^18 steps per launch keep launches few (measured on
    // the headline: 3.40e12 against 3.25e12 updates/s at 2^22 and 2^16).
    t.walk_chunk = 1u << 18;
    t.seg_candidates = 1u << 24;
    t.use_graph = 0;
    t.skip_table_check = 0;
    t.lds_segment = 15360;   // 256 threads x one 60-integer window
    t.walk_forms_off = 0;
    t.batch_walk_blocks = 0;
    return t;
}

extern "C" int pfc_device_open(int ordinal, pfc_device** out) {
    pfc::g_err[0] = 0;
    if (!out) return set_error(PFC_ERR_INVALID, "device_open: null out");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
    if (ordinal < 0 || ordinal >= n) return set_error(PFC_ERR_INVALID, "device ordinal out of range");
    pfc_device* d = new pfc_device();
    d->ordinal = ordinal;
    e = hipSetDevice(ordinal);
    if (e == hipSuccess) e = hipGetDeviceProperties(&d->prop, ordinal);
    if (e != hipSuccess) {
        delete d;
        return hip_fail(e, "hipGetDeviceProperties");
    }
    d->tuning = default_tuning();
    *out = d;
    return PFC_OK;
}

extern "C" int pfc_device_close(pfc_device* dev) {
    if (!dev) return PFC_OK;
    (void)hipSetDevice(dev->ordinal);
    pool_flush(dev);
    for (hipStream_t s : dev->pool.streams) (void)hipStreamDestroy(s);
    for (void* slab : dev->pool.slabs) (void)hipHostFree(slab);
    delete dev;
    return PFC_OK;
}

extern "C" int pfc_device_info(const pfc_device* dev, char* name, size_t cap, int* cus,
                               uint64_t* hbm) {
    if (!dev) return set_error(PFC_ERR_INVALID, "device_info: null device");
    if (name && cap) snprintf(name, cap, "%s", dev->prop.gcnArchName);
    if (cus) *cus = dev->prop.multiProcessorCount;
    if (hbm) *hbm = dev->prop.totalGlobalMem;
    return PFC_OK;
}

extern "C" int pfc_set_tuning(pfc_device* dev, const pfc_tuning* t) {
    if (!dev || !t) return set_error(PFC_ERR_INVALID, "set_tuning: null pointer");
    pfc_tuning d = default_tuning();
    pfc_tuning n = *t;
    if (!n.primes_per_lane) n.primes_per_lane = d.primes_per_lane;
    if (!n.block_size) n.block_size = d.block_size;
    if (!n.walk_chunk) n.walk_chunk = d.walk_chunk;
    if (!n.seg_candidates) n.seg_candidates = d.seg_candidates;
    if (n.primes_per_lane != 1 && n.primes_per_lane != 2 && n.primes_per_lane != 4)
        return set_error(PFC_ERR_INVALID, "primes_per_lane must be 1, 2 or 4");
    if (n.block_size != 64 && n.block_size != 128 && n.block_size != 256)
        return set_error(PFC_ERR_INVALID, "block_size must be 64, 128 or 256");
    if (n.seg_candidates > (1u << 28)) return set_error(PFC_ERR_CAPACITY, "seg_candidates > 2^28");
    if (!n.lds_segment) n.lds_segment = d.lds_segment;
    if (n.lds_segment != 15360 && n.lds_segment != 30720 && n.lds_segment != 61440)
        return set_error(PFC_ERR_INVALID, "lds_segment must be 15360, 30720 or 61440");
    if (n.walk_forms_off > 3) return set_error(PFC_ERR_INVALID, "walk_forms_off must be 0 to 3");
    if (n.batch_walk_blocks > pfc::kWalkMaxBlocks)
        return set_error(PFC_ERR_INVALID, "batch_walk_blocks must be 0 to 64");
    std::lock_guard<std::mutex> g(dev->mu);
    dev->tuning = n;
    return PFC_OK;
}

extern "C" int pfc_get_tuning(const pfc_device* dev, pfc_tuning* t) {
    if (!dev || !t) return set_error(PFC_ERR_INVALID, "get_tuning: null pointer");
    *t = tuning_of(dev);
    return PFC_OK;
}

extern "C" int pfc_device_malloc(pfc_device* dev, size_t bytes, void** ptr) {
    if (!dev || !ptr) return set_error(PFC_ERR_INVALID, "device_malloc: null pointer");
    {
        const hipError_t se = hipSetDevice(dev->ordinal);
        if (se != hipSuccess) return hip_fail(se, "device_malloc: hipSetDevice");
    }
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    return e == hipSuccess ? PFC_OK : hip_fail(e, "hipMalloc");
}

extern "C" int pfc_device_free(pfc_device* dev, void* ptr) {
    if (!dev) return set_error(PFC_ERR_INVALID, "device_free: null device");
    {
        const hipError_t se = hipSetDevice(dev->ordinal);
        if (se != hipSuccess) return hip_fail(se, "device_free: hipSetDevice");
    }
    hipError_t e = hipFree(ptr);
    return e == hipSuccess ? PFC_OK : hip_fail(e, "hipFree");
}

extern "C" int pfc_memcpy(pfc_device* dev, void* dst, const void* src, size_t bytes, int kind) {
    if (!dev || (!dst && bytes) || (!src && bytes)) return set_error(PFC_ERR_INVALID, "memcpy: null pointer");
    if (kind < 1 || kind > 3) return set_error(PFC_ERR_INVALID, "memcpy: kind must be 1, 2 or 3");
    {
        const hipError_t se = hipSetDevice(dev->ordinal);
        if (se != hipSuccess) return hip_fail(se, "memcpy: hipSetDevice");
    }
    const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : kind == 2 ? hipMemcpyDeviceToHost
                                                                         : hipMemcpyDeviceToDevice;
    hipError_t e = hipMemcpy(dst, src, bytes, k);
    return e == hipSuccess ? PFC_OK : hip_fail(e, "hipMemcpy");
}

extern "C" int pfc_device_synchronize(pfc_device* dev) {
    if (!dev) return set_error(PFC_ERR_INVALID, "synchronize: null device");
    {
        const hipError_t se = hipSetDevice(dev->ordinal);
        if (se != hipSuccess) return hip_fail(se, "synchronize: hipSetDevice");
    }
    hipError_t e = hipDeviceSynchronize();
    return e == hipSuccess ? PFC_OK : hip_fail(e, "hipDeviceSynchronize");
}

extern "C" int pfc_probe_copy(pfc_device* dev, uint64_t bytes, double* ms) {
    if (!dev) return set_error(PFC_ERR_INVALID, "probe_copy: null device");
    {
        const hipError_t se = hipSetDevice(dev->ordinal);
        if (se != hipSuccess) return hip_fail(se, "probe_copy: hipSetDevice");
    }
    return pfc::probe_copy(bytes, ms);
}

extern "C" int pfc_probe_mul_rates(pfc_device* dev, double rates[3]) {
    if (!dev || !rates) return set_error(PFC_ERR_INVALID, "probe: null pointer");
    {
        const hipError_t se = hipSetDevice(dev->ordinal);
        if (se != hipSuccess) return hip_fail(se, "probe: hipSetDevice");
    }
    return pfc::probe_mul_rates(dev->prop.multiProcessorCount, rates);
}

// ====================================================================== jobs

namespace {

// Validation of setupSearch (cl_sieve.cpp:600-641) plus the CLI ranges
// (main.cpp:58-72).
int validate(const pfc_job_desc* d) {
    if (!d) return set_error(PFC_ERR_INVALID, "null job description");
    const uint32_t m = d->modes;
    if (!(m & 7u) || (m & ~7u)) return set_error(PFC_ERR_INVALID, "-! or -# or -c argument is required");
    if ((m & PFC_MODE_PRIMORIAL) && (m & ~PFC_MODE_PRIMORIAL))
        return set_error(PFC_ERR_INVALID, "Select only one test type!");
    if (d->pmin == 0 || d->pmax == 0) return set_error(PFC_ERR_INVALID, "-p and -P arguments are required");
    if (d->nmin == 0 || d->nmax == 0) return set_error(PFC_ERR_INVALID, "-n and -N arguments are required");
    if (d->nmin > d->nmax) return set_error(PFC_ERR_INVALID, "nmin <= nmax is required");
    if (d->pmin > d->pmax) return set_error(PFC_ERR_INVALID, "pmin <= pmax is required");
    if (d->pmin < 3 || d->nmin < 101 || d->nmax > 0x7FFFFFFFu)
        return set_error(PFC_ERR_INVALID, "out of range: 3 <= p, 101 <= n, N < 2^31");
    return PFC_OK;
}

uint64_t default_fcap(const pfc_job_desc* d) {
    if (d->factor_capacity) return d->factor_capacity;
    return d->pmin < 0xFFFFFFFFull ? 30000000ull : 1000000ull;  // cl_sieve.cpp:645-647, main.cpp:215
}

// Montgomery-Vaughan: pi(x + y) - pi(x) <= 2y / ln y, plus margin for the
// base-2 pseudoprimes and the primes below 114.
uint64_t candidate_bound(uint64_t span) {
    if (span < 64) return 96;
    return (uint64_t)(2.1 * (double)span / log((double)span)) + 1024;
}

struct SegPlan {
    uint64_t span;  // integers per segment
    uint32_t cap;   // candidate capacity
};

SegPlan plan_segments(const pfc_job_desc* d, const pfc_tuning& t) {
    const double lnp = log((double)std::max<uint64_t>(d->pmax, 16));
    double span = (double)t.seg_candidates * lnp;
    const double total = (double)(d->pmax - d->pmin);
    if (span > total) span = total;
    if (span < 1) span = 1;
    SegPlan s;
    s.span = (uint64_t)span;
    uint64_t c = candidate_bound(s.span);
    if (c > 0xF0000000ull) c = 0xF0000000ull;
    s.cap = (uint32_t)c;
    return s;
}

}  // namespace

struct pfc_job {
    pfc_device* dev = nullptr;
    pfc_job_desc desc{};
    pfc_tuning tuning{};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evw0 = nullptr, evw1 = nullptr;
    SegPlan plan{};
    uint64_t fcap = 0;
    uint64_t pos = 0;
    // host tables
    HostTable ta, tb;
    std::vector<uint32_t> list;       // primorial multipliers
    std::vector<uint32_t> composite;  // bitmap over n - nmin (C and F|C)
    std::vector<uint32_t> clist;      // the composites in [nmin, nmax), increasing (C and F|C)
    std::vector<uint32_t> cprefix;    // count of clist entries before each bitmap word
    uint32_t jbegin = 0, jend = 0;    // walk step range
    uint64_t steps_per_cand = 0;      // multiplier applications per candidate over the walk
    uint32_t debug_short = 0;         // test hook: steps left out of each segment's last walk launch
    // device memory
    uint64_t *d_ca = nullptr, *d_cb = nullptr;
    uint32_t *d_ga = nullptr, *d_gb = nullptr;
    uint32_t* d_list = nullptr;
    uint32_t* d_comp = nullptr;
    uint32_t *d_clist = nullptr, *d_cprefix = nullptr;
    uint64_t* d_p = nullptr;
    uint32_t* d_v = nullptr;
    uint64_t *d_one = nullptr, *d_ra = nullptr, *d_rb = nullptr;
    pfc_ctrl* d_ctrl = nullptr;
    pfc_factor* d_fac = nullptr;
    pfc_ctrl* h_ctrl = nullptr;  // pinned
    // start-table self-check, queued at creation and read at the first drain
    uint64_t* d_verify = nullptr;   // 2 x PFC_VERIFY_WORDS
    uint64_t* h_verify = nullptr;   // pinned: {product A, table A, product B, table B}
    int verify_tables_n = 0;        // tables still to be checked (0: none pending)
    // captured segment pipeline (tuning.use_graph)
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    uint64_t* d_bounds = nullptr;   // {lo, hi} read by the captured kernels
    uint64_t* h_bounds = nullptr;   // pinned
    // results
    uint64_t dev_checksum = 0, primecount = 0, segments = 0;
    double device_ms = 0, walk_ms = 0;
    uint64_t walk_launches = 0, walk_updates = 0;
    std::vector<pfc_factor> factors;
    std::vector<std::pair<void*, size_t>> allocs;   // device pool buffers and their size classes
};

namespace {

// A device buffer of the job from the device pool, returned at destruction.
template <class T>
int job_alloc(pfc_job* j, T** p, size_t bytes) {
    size_t cls = 0;
    int rc = pool_malloc(j->dev, (void**)p, bytes, &cls);
    if (!rc) j->allocs.emplace_back((void*)*p, cls);
    return rc;
}

template <class T>
int job_pinned(pfc_job* j, T** p, size_t bytes) {
    if (bytes > kPinnedBlock) return set_error(PFC_ERR_INVALID, "pinned block too small");
    return pool_pinned(j->dev, (void**)p);
}

// Everything a job holds goes back to the device pool once its stream has
// drained (nothing queued may still use a buffer that another job reuses).
void free_job_device(pfc_job* j) {
    if (j->stream) (void)hipStreamSynchronize(j->stream);
    for (auto& a : j->allocs) pool_free(j->dev, a.first, a.second);
    j->allocs.clear();
    pool_pinned_put(j->dev, j->h_ctrl);
    pool_pinned_put(j->dev, j->h_verify);
    pool_pinned_put(j->dev, j->h_bounds);
    if (j->gexec) (void)hipGraphExecDestroy(j->gexec);
    if (j->graph) (void)hipGraphDestroy(j->graph);
    if (j->ev0) (void)hipEventDestroy(j->ev0);
    if (j->ev1) (void)hipEventDestroy(j->ev1);
    if (j->evw0) (void)hipEventDestroy(j->evw0);
    if (j->evw1) (void)hipEventDestroy(j->evw1);
    if (j->stream) pool_stream_put(j->dev, j->stream);
}

template <class T>
int upload(pfc_job* j, T** dst, const std::vector<T>& src, const char* what) {
    const size_t bytes = std::max<size_t>(src.size(), 1) * sizeof(T);
    int rc = job_alloc(j, dst, bytes);
    if (rc) return rc;
    hipStream_t st = j->stream;
    hipError_t e = hipSuccess;
    if (!src.empty()) {
        e = hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return hip_fail(e, what);
    }
    return PFC_OK;
}

// Step range and multiplier data of the walk (cl_sieve.cpp:1478-1499,
// 996-1000, 1160-1170): F and C walk n in [nmin, nmax); P walks the primes in
// [nmin, nmax-1] and requires nmax composite (cl_sieve.cpp:1058-1066).
int build_walk_data(pfc_job* j) {
    const pfc_job_desc& d = j->desc;
    const uint32_t N = d.nmin - 1;
    if (d.modes == PFC_MODE_PRIMORIAL) {
        j->ta = build_table(1, N);
        j->list = primes_in(d.nmin, d.nmax - 1);
        std::vector<uint32_t> last = primes_in(d.nmax, d.nmax);
        if (!last.empty()) return set_error(PFC_ERR_SIEVE, "CPU sieve failure.");
        j->jbegin = 0;
        j->jend = (uint32_t)j->list.size();
        j->steps_per_cand = j->list.size();
        return PFC_OK;
    }
    j->jbegin = d.nmin;
    j->jend = d.nmax;
    j->steps_per_cand = 0;
    if (d.modes & PFC_MODE_FACTORIAL) j->ta = build_table(0, N);
    if (d.modes & PFC_MODE_COMPOSITORIAL) {
        if (d.modes & PFC_MODE_FACTORIAL) j->tb = build_table(2, N);
        else j->ta = build_table(2, N);
        const uint32_t w = d.nmax - d.nmin;
        j->composite.assign(w / 32 + 2, 0xFFFFFFFFu);
        const std::vector<uint32_t> pr = primes_in(d.nmin, d.nmax - 1);
        for (uint32_t q : pr) {
            const uint32_t off = q - d.nmin;
            j->composite[off >> 5] &= ~(1u << (off & 31));
        }
        j->steps_per_cand += w - pr.size();
        // The walk reads the composites 16 at a time from this list; the
        // prefix counts let it find a launch's part of the list on the device.
        if (w <= kMaxCompositeList) {
            j->clist.reserve(w - pr.size());
            j->cprefix.assign(j->composite.size(), 0);
            for (uint32_t k = 0; k < j->composite.size(); ++k) {
                j->cprefix[k] = (uint32_t)j->clist.size();
                uint32_t word = j->composite[k];
                if ((uint64_t)k * 32 + 32 > w) word &= k * 32 >= w ? 0u : (1u << (w - k * 32)) - 1u;
                for (; word; word &= word - 1) j->clist.push_back(d.nmin + k * 32 + (uint32_t)__builtin_ctz(word));
            }
        }
    }
    if (d.modes & PFC_MODE_FACTORIAL) j->steps_per_cand += d.nmax - d.nmin;
    return PFC_OK;
}

uint64_t footprint_of(const pfc_job_desc* d, const pfc_tuning& t) {
    const SegPlan s = plan_segments(d, t);
    const bool two = (d->modes & PFC_MODE_FACTORIAL) && (d->modes & PFC_MODE_COMPOSITORIAL);
    uint64_t per = 8 + 4 + 8 + 8 + (two ? 8 : 0);
    uint64_t bytes = (uint64_t)s.cap * per + default_fcap(d) * sizeof(pfc_factor) + sizeof(pfc_ctrl);
    const uint64_t w = (uint64_t)(d->nmax - d->nmin);
    if (d->modes & PFC_MODE_COMPOSITORIAL) bytes += w / 8 + 8 + (w <= kMaxCompositeList ? 4 * w + w / 8 + 8 : 0);
    if (d->modes & PFC_MODE_PRIMORIAL) bytes += 4 * (uint64_t)(1.3 * (double)w / log((double)d->nmax) + 64);
    // start tables: about one 32-bit chunk per prime below nmin
    bytes += 8 * (uint64_t)(1.3 * (double)d->nmin / log((double)d->nmin) + 64);
    return bytes;
}

// Device self-check of the job's start tables against the product of their
// multipliers modulo 2^64-59 (cl_sieve.cpp:917-970, 1079-1134, 1246-1304).
// Queued on the job's stream at creation; the two results per table land in
// pinned memory and are compared by check_tables when the job first
// synchronises, so the check overlaps the host work that follows creation.
int queue_table_check(pfc_job* j) {
    const bool two = (j->desc.modes & PFC_MODE_FACTORIAL) && (j->desc.modes & PFC_MODE_COMPOSITORIAL);
    int rc0 = job_alloc(j, &j->d_verify, 2 * PFC_VERIFY_WORDS * sizeof(uint64_t));
    if (!rc0) rc0 = job_pinned(j, &j->h_verify, 4 * sizeof(uint64_t));
    if (rc0) return rc0;
    hipError_t e = hipSuccess;
    const HostTable* hs[2] = {&j->ta, &j->tb};
    const uint64_t* cs[2] = {j->d_ca, j->d_cb};
    const uint32_t* gs[2] = {j->d_ga, j->d_gb};
    const int n = two ? 2 : 1;
    for (int i = 0; i < n; ++i) {
        const pfc_setup_table t{cs[i], gs[i], hs[i]->npos, hs[i]->digit_bits};
        uint64_t* work = j->d_verify + (size_t)i * PFC_VERIFY_WORDS;
        int rc = pfc::launch_verify_table(hs[i]->kind, hs[i]->N, &t, work, work, j->stream);
        if (rc) return rc;
        e = hipMemcpyAsync(j->h_verify + 2 * i, work, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, j->stream);
        if (e != hipSuccess) return hip_fail(e, "table check");
    }
    j->verify_tables_n = n;
    return PFC_OK;
}

// Compares the queued self-check results (waits for them if needed).
int check_tables(pfc_job* j) {
    if (!j->verify_tables_n) return PFC_OK;
    hipError_t e = hipStreamSynchronize(j->stream);
    if (e != hipSuccess) return hip_fail(e, "table check");
    const HostTable* hs[2] = {&j->ta, &j->tb};
    for (int i = 0; i < j->verify_tables_n; ++i)
        if (j->h_verify[2 * i] != j->h_verify[2 * i + 1])
            return set_error(PFC_ERR_VALIDATION, hs[i]->kind == 0 ? "error: power table verification failed"
                                                                  : "error: product/prime table verification failed");
    j->verify_tables_n = 0;
    return PFC_OK;
}

}  // namespace

extern "C" int pfc_job_footprint(const pfc_device* dev, const pfc_job_desc* desc, uint64_t* bytes) {
    pfc::g_err[0] = 0;
    int rc = validate(desc);
    if (rc) return rc;
    if (!bytes) return set_error(PFC_ERR_INVALID, "footprint: null out");
    *bytes = footprint_of(desc, dev ? tuning_of(dev) : default_tuning());
    return PFC_OK;
}

extern "C" int pfc_job_create(pfc_device* dev, const pfc_job_desc* desc, pfc_job** out) {
    pfc::g_err[0] = 0;
    if (!dev || !out) return set_error(PFC_ERR_INVALID, "job_create: null pointer");
    int rc = validate(desc);
    if (rc) return rc;
    {
        const hipError_t se = hipSetDevice(dev->ordinal);
        if (se != hipSuccess) return hip_fail(se, "job_create: hipSetDevice");
    }
    pfc_job* j = new pfc_job();
    j->dev = dev;
    j->desc = *desc;
    {
        std::lock_guard<std::mutex> g(dev->mu);
        j->tuning = dev->tuning;
    }
    j->plan = plan_segments(desc, j->tuning);
    j->fcap = default_fcap(desc);
    if (j->fcap > 0xFFFFFFFFull) j->fcap = 0xFFFFFFFFull;
    j->pos = desc->pmin;
    rc = build_walk_data(j);
    if (rc) {
        delete j;
        return rc;
    }
    rc = pool_stream(dev, &j->stream);
    if (rc) {
        delete j;
        return rc;
    }
    hipError_t e = hipEventCreate(&j->ev0);
    if (e == hipSuccess) e = hipEventCreate(&j->ev1);
    if (e == hipSuccess) e = hipEventCreate(&j->evw0);
    if (e == hipSuccess) e = hipEventCreate(&j->evw1);
    if (e != hipSuccess) {
        rc = hip_fail(e, "stream/event create");
        free_job_device(j);
        delete j;
        return rc;
    }
    const uint32_t cap = j->plan.cap;
    const bool two = (desc->modes & PFC_MODE_FACTORIAL) && (desc->modes & PFC_MODE_COMPOSITORIAL);
    rc = job_alloc(j, &j->d_p, (size_t)cap * 8);
    if (!rc) rc = job_alloc(j, &j->d_v, (size_t)cap * 4);
    if (!rc) rc = job_alloc(j, &j->d_one, (size_t)cap * 8);
    if (!rc) rc = job_alloc(j, &j->d_ra, (size_t)cap * 8);
    if (!rc) rc = job_alloc(j, &j->d_ctrl, sizeof(pfc_ctrl));
    if (!rc) rc = job_alloc(j, &j->d_fac, (size_t)j->fcap * sizeof(pfc_factor));
    if (!rc && two) rc = job_alloc(j, &j->d_rb, (size_t)cap * 8);
    if (!rc) rc = job_pinned(j, &j->h_ctrl, sizeof(pfc_ctrl));
    if (!rc) rc = upload(j, &j->d_ca, j->ta.chunks, "table A");
    if (!rc) rc = upload(j, &j->d_ga, j->ta.groups, "table A groups");
    if (!rc && two) rc = upload(j, &j->d_cb, j->tb.chunks, "table B");
    if (!rc && two) rc = upload(j, &j->d_gb, j->tb.groups, "table B groups");
    if (!rc && !j->list.empty()) rc = upload(j, &j->d_list, j->list, "multiplier list");
    if (!rc && !j->composite.empty()) rc = upload(j, &j->d_comp, j->composite, "composite map");
    if (!rc && !j->cprefix.empty()) rc = upload(j, &j->d_clist, j->clist, "composite list");
    if (!rc && !j->cprefix.empty()) rc = upload(j, &j->d_cprefix, j->cprefix, "composite counts");
    if (!rc) {
        e = hipStreamSynchronize(j->stream);
        if (e != hipSuccess) rc = hip_fail(e, "table upload");
    }
    if (!rc && !j->tuning.skip_table_check) rc = queue_table_check(j);
    if (rc) {
        // free_job_device waits for the stream first: copies queued from the
        // host vectors may still be reading them.
        free_job_device(j);
        delete j;
        return rc;
    }
    // Only the device copies of the composite data are read from here on (the
    // list alone can reach 512 MB); the stream has drained the uploads.
    std::vector<uint32_t>().swap(j->composite);
    std::vector<uint32_t>().swap(j->clist);
    std::vector<uint32_t>().swap(j->cprefix);
    *out = j;
    return PFC_OK;
}

// Drains the control block and factor records of a finished segment.
static int drain_segment(pfc_job* j, bool* overflow, uint64_t launches, bool timed);

// Message of a stored reciprocal failing its check at the fold.
static const char* const kReciprocalFailure = "error: gpu state check failure (stored reciprocal)";

// Walk-completion check (the reference's check.cl:62-73 guard, reported as
// cl_sieve.cpp:476-480 does): every candidate of a finished segment must have
// taken every multiplier step of the job's walk.
static int check_walked(const pfc_job* j, const pfc_ctrl& c) {
    const uint64_t want = (uint64_t)c.ncand * j->steps_per_cand;
    if (c.walked == want) return PFC_OK;
    char msg[160];
    snprintf(msg, sizeof(msg), "error: gpu validation failure (walked %llu of %llu multiplier steps)",
             (unsigned long long)c.walked, (unsigned long long)want);
    return set_error(PFC_ERR_VALIDATION, msg);
}

// Captures clear -> primes below 114 -> generate -> setup -> walk chunks ->
// check for a job as one HIP graph; segment bounds come from device memory.
// Largest multiplier of a job's walk steps below b: n = j, or in P mode the
// j-th prime of the ascending list.  The walk pairs its steps (tests every
// second one) on wavefronts whose p all exceed twice this bound.
// End of a walk launch that would end at b: the job's last step, less the
// steps the completion-check test hook leaves out (pfc_job_debug_shorten_walk).
static uint32_t launch_end(const pfc_job* j, uint32_t a, uint32_t b) {
    if (b != j->jend || j->debug_short == 0) return b;
    return b - a > j->debug_short ? b - j->debug_short : a;
}

static uint32_t walk_bound(const pfc_job* j, uint64_t b) {
    if (b == 0) return 0;
    if (j->desc.modes == PFC_MODE_PRIMORIAL) return b <= j->list.size() ? j->list[b - 1] : 0;
    return (uint32_t)(b - 1);
}

static int build_graph(pfc_job* j) {
    hipStream_t st = j->stream;
    const uint32_t cap = j->plan.cap;
    const pfc_job_desc& d = j->desc;
    int rc0 = job_alloc(j, &j->d_bounds, 2 * sizeof(uint64_t));
    if (!rc0) rc0 = job_pinned(j, &j->h_bounds, 2 * sizeof(uint64_t));
    if (rc0) return rc0;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
    int rc = pfc::launch_head_dev(j->d_bounds, j->plan.span, j->d_p, cap, j->d_ctrl, st, j->tuning.lds_segment);
    pfc_setup_table ta{j->d_ca, j->d_ga, j->ta.npos, j->ta.digit_bits};
    pfc_setup_table tb{j->d_cb, j->d_gb, j->tb.npos, j->tb.digit_bits};
    if (!rc)
        rc = pfc::launch_setup(j->d_p, j->d_ctrl, j->d_v, j->d_one, j->d_ra, j->d_rb, &ta,
                               j->d_rb ? &tb : nullptr, cap, st);
    const uint32_t chunk = std::max<uint32_t>(j->tuning.walk_chunk, 1);
    for (uint64_t a = j->jbegin; a < j->jend && !rc; a += chunk) {
        const uint32_t b = launch_end(j, (uint32_t)a, (uint32_t)std::min<uint64_t>(j->jend, a + chunk));
        rc = pfc::launch_walk(d.modes, (uint32_t)a, b, j->d_list, j->d_comp, j->d_clist, j->d_cprefix, d.nmin,
                              j->d_p, j->d_v, j->d_ra, j->d_rb, j->d_ctrl, j->d_fac, (uint32_t)j->fcap,
                              cap, j->tuning.primes_per_lane, j->tuning.block_size, walk_bound(j, b),
                              j->tuning.walk_forms_off, st);
    }
    if (!rc)
        rc = pfc::launch_check(d.modes, d.nmax, j->d_p, j->d_v, j->d_one, j->d_ra, j->d_rb, j->d_ctrl,
                               cap, st);
    e = hipStreamEndCapture(st, &j->graph);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&j->gexec, j->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    return PFC_OK;
}

static int run_segment_graph(pfc_job* j, uint64_t lo, uint64_t hi, bool* overflow) {
    if (!j->gexec) {
        int rc = build_graph(j);
        if (rc) return rc;
    }
    hipStream_t st = j->stream;
    j->h_bounds[0] = lo;
    j->h_bounds[1] = hi;
    hipError_t e = hipMemcpyAsync(j->d_bounds, j->h_bounds, 2 * sizeof(uint64_t), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipGraphLaunch(j->gexec, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(j->h_ctrl, j->d_ctrl, sizeof(pfc_ctrl), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "graph segment");
    return drain_segment(j, overflow, 0, false);
}

// One p-segment [lo, hi): generate, setup, walk in chunks, check, then drain
// the control block and factor records (cl_sieve.cpp:1563-1726, 450-494).
static int run_segment(pfc_job* j, uint64_t lo, uint64_t hi, bool* overflow) {
    if (j->tuning.use_graph) return run_segment_graph(j, lo, hi, overflow);
    hipStream_t st = j->stream;
    const uint32_t cap = j->plan.cap;
    const pfc_job_desc& d = j->desc;
    *overflow = false;
    hipError_t e = hipMemsetAsync(j->d_ctrl, 0, sizeof(pfc_ctrl), st);
    if (e != hipSuccess) return hip_fail(e, "ctrl clear");
    int rc = pfc::launch_generate(lo, hi, j->d_p, cap, j->d_ctrl, st, j->tuning.lds_segment);
    if (rc) return rc;
    pfc_setup_table ta{j->d_ca, j->d_ga, j->ta.npos, j->ta.digit_bits};
    pfc_setup_table tb{j->d_cb, j->d_gb, j->tb.npos, j->tb.digit_bits};
    rc = pfc::launch_setup(j->d_p, j->d_ctrl, j->d_v, j->d_one, j->d_ra, j->d_rb, &ta,
                           j->d_rb ? &tb : nullptr, cap, st);
    if (rc) return rc;
    const uint32_t chunk = std::max<uint32_t>(j->tuning.walk_chunk, 1);
    uint64_t launches = 0;
    e = hipEventRecord(j->evw0, st);
    if (e != hipSuccess) return hip_fail(e, "walk timing");
    for (uint64_t a = j->jbegin; a < j->jend; a += chunk, ++launches) {
        const uint32_t b = launch_end(j, (uint32_t)a, (uint32_t)std::min<uint64_t>(j->jend, a + chunk));
        rc = pfc::launch_walk(d.modes, (uint32_t)a, b, j->d_list, j->d_comp, j->d_clist, j->d_cprefix, d.nmin,
                              j->d_p, j->d_v, j->d_ra, j->d_rb, j->d_ctrl, j->d_fac, (uint32_t)j->fcap,
                              cap, j->tuning.primes_per_lane, j->tuning.block_size, walk_bound(j, b),
                              j->tuning.walk_forms_off, st);
        if (rc) return rc;
    }
    e = hipEventRecord(j->evw1, st);
    if (e != hipSuccess) return hip_fail(e, "walk timing");
    rc = pfc::launch_check(d.modes, d.nmax, j->d_p, j->d_v, j->d_one, j->d_ra, j->d_rb, j->d_ctrl,
                           cap, st);
    if (rc) return rc;
    e = hipMemcpyAsync(j->h_ctrl, j->d_ctrl, sizeof(pfc_ctrl), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "segment");
    return drain_segment(j, overflow, launches, true);
}

static int drain_segment(pfc_job* j, bool* overflow, uint64_t launches, bool timed) {
    *overflow = false;
    int rc = check_tables(j);
    if (rc) return rc;
    hipError_t e;
    const pfc_ctrl c = *j->h_ctrl;
    if (c.flags & 1u) {
        *overflow = true;
        return PFC_OK;
    }
    if (c.flags & 2u) return set_error(PFC_ERR_VALIDATION, kReciprocalFailure);
    if (c.nfactor > j->fcap) {
        // More factor records than the buffer holds (the reference exits here,
        // cl_sieve.cpp:484-487, but its segments are far smaller): the segment
        // is redone in halves, like a candidate overflow.
        *overflow = true;
        return PFC_OK;
    }
    rc = check_walked(j, c);
    if (rc) return rc;
    if (c.nfactor) {
        const size_t old = j->factors.size();
        j->factors.resize(old + c.nfactor);
        e = hipMemcpy(j->factors.data() + old, j->d_fac, (size_t)c.nfactor * sizeof(pfc_factor),
                      hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e, "factor copy");
    }
    j->dev_checksum += c.checksum;
    j->primecount += c.primecount;
    j->segments += 1;
    if (timed) {
        float wms = 0;
        e = hipEventElapsedTime(&wms, j->evw0, j->evw1);
        if (e != hipSuccess) return hip_fail(e, "walk timing");
        j->walk_ms += wms;
        j->walk_launches += launches;
        j->walk_updates += (uint64_t)c.ncand * j->steps_per_cand;
    }
    return PFC_OK;
}

extern "C" int pfc_job_run(pfc_job* j, uint64_t max_segments) {
    pfc::g_err[0] = 0;
    if (!j) return set_error(PFC_ERR_INVALID, "job_run: null job");
    hipError_t e = hipSetDevice(j->dev->ordinal);
    if (e == hipSuccess) e = hipEventRecord(j->ev0, j->stream);
    if (e != hipSuccess) return hip_fail(e, "job_run start");
    uint64_t done = 0;
    uint64_t span = j->plan.span;
    while (j->pos < j->desc.pmax && (max_segments == 0 || done < max_segments)) {
        uint64_t hi = j->pos + span;
        if (hi > j->desc.pmax || hi < j->pos) hi = j->desc.pmax;
        bool overflow = false;
        int rc = run_segment(j, j->pos, hi, &overflow);
        if (rc) return rc;
        if (overflow) {  // more candidates or factor records than the buffers hold: split the segment
            if (span <= 1) {
                const pfc_ctrl& c = *j->h_ctrl;
                if (c.nfactor > j->fcap) {
                    char msg[128];
                    snprintf(msg, sizeof(msg), "Error: number of results (%u) overflowed array.", c.nfactor);
                    return set_error(PFC_ERR_FACTOR_OVERFLOW, msg);
                }
                return set_error(PFC_ERR_CAPACITY, "candidate buffer overflow");
            }
            span /= 2;
            continue;
        }
        j->pos = hi;
        ++done;
    }
    e = hipEventRecord(j->ev1, j->stream);
    if (e == hipSuccess) e = hipEventSynchronize(j->ev1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, j->ev0, j->ev1);
    if (e != hipSuccess) return hip_fail(e, "job_run end");
    j->device_ms += ms;
    return PFC_OK;
}

extern "C" int pfc_job_position(const pfc_job* j, uint64_t* p) {
    if (!j || !p) return set_error(PFC_ERR_INVALID, "job_position: null pointer");
    *p = j->pos;
    return PFC_OK;
}

extern "C" int pfc_job_seek(pfc_job* j, uint64_t p, uint64_t dev_checksum, uint64_t primecount) {
    if (!j) return set_error(PFC_ERR_INVALID, "job_seek: null job");
    if (p < j->desc.pmin || p > j->desc.pmax) return set_error(PFC_ERR_INVALID, "seek outside [pmin, pmax]");
    j->pos = p;
    j->dev_checksum = dev_checksum;
    j->primecount = primecount;
    return PFC_OK;
}

extern "C" int pfc_job_stats_get(const pfc_job* j, pfc_job_stats* s) {
    if (!j || !s) return set_error(PFC_ERR_INVALID, "job_stats: null pointer");
    s->dev_checksum = j->dev_checksum;
    s->primecount = j->primecount;
    s->raw_factors = j->factors.size();
    s->segments = j->segments;
    s->device_ms = j->device_ms;
    s->walk_ms = j->walk_ms;
    s->walk_launches = j->walk_launches;
    s->walk_updates = j->walk_updates;
    return PFC_OK;
}

static bool factor_less(const pfc_factor& x, const pfc_factor& y) {
    if (x.p != y.p) return x.p < y.p;
    const int32_t nx = x.nc < 0 ? -x.nc : x.nc, ny = y.nc < 0 ? -y.nc : y.nc;
    if (nx != ny) return nx < ny;
    return x.type < y.type;
}

extern "C" int pfc_job_factors(const pfc_job* j, pfc_factor* out, size_t cap, size_t* n) {
    if (!j || !n) return set_error(PFC_ERR_INVALID, "job_factors: null pointer");
    *n = j->factors.size();
    if (!out) return PFC_OK;
    std::vector<pfc_factor> f = j->factors;
    std::sort(f.begin(), f.end(), factor_less);
    const size_t k = std::min(cap, f.size());
    if (k) memcpy(out, f.data(), k * sizeof(pfc_factor));
    return PFC_OK;
}

extern "C" int pfc_job_sync(pfc_job* j) {
    pfc::g_err[0] = 0;
    if (!j) return set_error(PFC_ERR_INVALID, "job_sync: null job");
    hipError_t e = hipSetDevice(j->dev->ordinal);
    if (e == hipSuccess) e = hipStreamSynchronize(j->stream);
    if (e != hipSuccess) return hip_fail(e, "job_sync");
    return check_tables(j);
}

extern "C" int pfc_job_result(const pfc_job* j, uint64_t* checksum_partial, uint64_t* primecount,
                              pfc_factor* out, size_t cap, size_t* n) {
    if (!j || !checksum_partial || !primecount || !n) return set_error(PFC_ERR_INVALID, "job_result: null pointer");
    *checksum_partial = j->dev_checksum;
    *primecount = j->primecount;
    return pfc_job_factors(j, out, cap, n);
}

extern "C" int pfc_job_clear_factors(pfc_job* j) {
    if (!j) return set_error(PFC_ERR_INVALID, "job_clear_factors: null job");
    j->factors.clear();
    return PFC_OK;
}

// The job is released even when the device reports an error (its buffers go
// back to the pool only after its stream has drained); the error is returned.
extern "C" int pfc_job_destroy(pfc_job* j) {
    if (!j) return PFC_OK;
    pfc::g_err[0] = 0;
    hipError_t e = hipSetDevice(j->dev->ordinal);
    if (e == hipSuccess && j->stream) e = hipStreamSynchronize(j->stream);
    free_job_device(j);
    delete j;
    return e == hipSuccess ? PFC_OK : hip_fail(e, "job_destroy");
}

extern "C" int pfc_job_debug_shorten_walk(pfc_job* j, uint32_t steps) {
    if (!j) return set_error(PFC_ERR_INVALID, "debug_shorten_walk: null job");
    j->debug_short = steps;
    if (j->gexec) (void)hipGraphExecDestroy(j->gexec);
    if (j->graph) (void)hipGraphDestroy(j->graph);
    j->gexec = nullptr;
    j->graph = nullptr;
    return PFC_OK;
}

// Times walk launches of every (primes_per_lane, block_size) shape on a
// sample of the job (a p-window with about 2^20 candidates, generated and set
// up once) and keeps the fastest in the device tuning.  The counterpart of
// the reference's profile-driven sizing (cl_sieve.cpp:669-769, 1642-1694).
extern "C" int pfc_autotune(pfc_device* dev, const pfc_job_desc* desc, pfc_tuning* best) {
    pfc::g_err[0] = 0;
    if (!dev || !desc) return set_error(PFC_ERR_INVALID, "autotune: null pointer");
    int rc = validate(desc);
    if (rc) return rc;
    pfc_job_desc d = *desc;
    const double lnp = log((double)std::max<uint64_t>(d.pmin, 16));
    const uint64_t want = (uint64_t)((double)(1u << 20) * lnp);
    if (d.pmax - d.pmin > want) d.pmax = d.pmin + want;
    pfc_tuning saved = tuning_of(dev);
    pfc_tuning t = saved;
    t.seg_candidates = 1u << 20;
    {
        std::lock_guard<std::mutex> g(dev->mu);
        dev->tuning = t;
    }
    pfc_job* j = nullptr;
    rc = pfc_job_create(dev, &d, &j);
    if (rc) {
        std::lock_guard<std::mutex> g(dev->mu);
        dev->tuning = saved;
        return rc;
    }
    hipStream_t st = j->stream;
    const uint32_t cap = j->plan.cap;
    hipError_t e = hipMemsetAsync(j->d_ctrl, 0, sizeof(pfc_ctrl), st);
    if (e != hipSuccess) rc = hip_fail(e, "autotune clear");
    if (!rc) rc = pfc::launch_generate(d.pmin, d.pmin + j->plan.span, j->d_p, cap, j->d_ctrl, st, t.lds_segment);
    pfc_setup_table ta{j->d_ca, j->d_ga, j->ta.npos, j->ta.digit_bits};
    pfc_setup_table tb{j->d_cb, j->d_gb, j->tb.npos, j->tb.digit_bits};
    if (!rc)
        rc = pfc::launch_setup(j->d_p, j->d_ctrl, j->d_v, j->d_one, j->d_ra, j->d_rb, &ta,
                               j->d_rb ? &tb : nullptr, cap, st);
    const uint32_t steps = std::min<uint32_t>(4096, j->jend - j->jbegin);
    float best_ms = 1e30f;
    pfc_tuning win = saved;
    const uint32_t ppls[3] = {1, 2, 4}, blocks[3] = {64, 128, 256};
    for (int rep = 0; rep < 2 && !rc; ++rep) {      // first pass warms up
        for (uint32_t ppl : ppls) {
            for (uint32_t blk : blocks) {
                if (rc) break;
                (void)hipEventRecord(j->ev0, st);
                rc = pfc::launch_walk(d.modes, j->jbegin, j->jbegin + steps, j->d_list, j->d_comp, j->d_clist,
                                      j->d_cprefix, d.nmin, j->d_p, j->d_v, j->d_ra, j->d_rb, j->d_ctrl, j->d_fac,
                                      (uint32_t)j->fcap, cap, ppl, blk, walk_bound(j, j->jbegin + steps),
                                      j->tuning.walk_forms_off, st);
                (void)hipEventRecord(j->ev1, st);
                e = hipEventSynchronize(j->ev1);
                if (e != hipSuccess) rc = hip_fail(e, "autotune walk");
                float ms = 0;
                (void)hipEventElapsedTime(&ms, j->ev0, j->ev1);
                if (rep == 1 && ms < best_ms) {
                    best_ms = ms;
                    win.primes_per_lane = ppl;
                    win.block_size = blk;
                }
            }
        }
    }
    // n-chunk length (multiplier steps per walk launch) at the chosen shape:
    // the same step total walked in launches of each length.
    const uint32_t total = std::min<uint32_t>(1u << 18, j->jend - j->jbegin);
    const uint32_t chunks[3] = {1u << 14, 1u << 16, 1u << 18};
    float best_chunk_ms = 1e30f;
    for (int rep = 0; rep < 2 && !rc; ++rep) {
        for (uint32_t c : chunks) {
            if (rc || (c > total && c != chunks[0])) break;
            (void)hipEventRecord(j->ev0, st);
            for (uint32_t a = 0; a < total && !rc; a += c)
                rc = pfc::launch_walk(d.modes, j->jbegin + a, j->jbegin + std::min(total, a + c), j->d_list,
                                      j->d_comp, j->d_clist, j->d_cprefix, d.nmin, j->d_p, j->d_v, j->d_ra, j->d_rb, j->d_ctrl,
                                      j->d_fac, (uint32_t)j->fcap, cap, win.primes_per_lane, win.block_size,
                                      walk_bound(j, j->jbegin + std::min(total, a + c)),
                                      j->tuning.walk_forms_off, st);
            (void)hipEventRecord(j->ev1, st);
            e = hipEventSynchronize(j->ev1);
            if (e != hipSuccess && !rc) rc = hip_fail(e, "autotune walk chunk");
            float ms = 0;
            (void)hipEventElapsedTime(&ms, j->ev0, j->ev1);
            if (rep == 1 && ms < best_chunk_ms) {
                best_chunk_ms = ms;
                win.walk_chunk = c;
            }
        }
    }
    // LDS segment of the generator: the sample window generated at each size.
    const uint32_t segs[3] = {15360, 30720, 61440};
    float best_gen_ms = 1e30f;
    for (int rep = 0; rep < 2 && !rc; ++rep) {
        for (uint32_t lds : segs) {
            if (rc) break;
            e = hipMemsetAsync(j->d_ctrl, 0, sizeof(pfc_ctrl), st);
            if (e != hipSuccess) rc = hip_fail(e, "autotune clear");
            (void)hipEventRecord(j->ev0, st);
            if (!rc) rc = pfc::launch_generate(d.pmin, d.pmin + j->plan.span, j->d_p, cap, j->d_ctrl, st, lds);
            (void)hipEventRecord(j->ev1, st);
            e = hipEventSynchronize(j->ev1);
            if (e != hipSuccess && !rc) rc = hip_fail(e, "autotune generate");
            float ms = 0;
            (void)hipEventElapsedTime(&ms, j->ev0, j->ev1);
            if (rep == 1 && ms < best_gen_ms) {
                best_gen_ms = ms;
                win.lds_segment = lds;
            }
        }
    }
    pfc_job_destroy(j);
    {
        std::lock_guard<std::mutex> g(dev->mu);
        dev->tuning = rc ? saved : win;
    }
    if (best) *best = tuning_of(dev);
    return rc;
}

// ====================================================================== batcher
//
// Continuous batching of many jobs on one device.  Every round:
//   1. jobs that need a new p-segment get one (clear control block, generate
//      candidates) -- all jobs' generation launches are queued back to back;
//   2. the candidates of every active segment are cut into work items of
//      blocks of 64*K candidates x up to `max_steps` multiplier steps (or a
//      segment's share of a round sized by work), described by one range per
//      segment and ordered by decreasing cost (longest first), and
//      persistent kernels drain them from device queues (a counter each;
//      an item finds its segment by binary search): the set-up of segments
//      that start in this round, the walk of each walk mode, the checksum
//      fold of segments whose walk ends in this round;
//   3. finished segments are drained to the host like pfc_job_run does;
//      finished jobs are returned to the caller.
// Jobs can be added between rounds; their state stays resident on the device.

// Work items per batch launch at most (the queue counter is 32-bit); segments
// beyond it wait for the next round.
constexpr uint64_t kMaxBatchItems = 1ull << 31;

struct pfc_batcher {
    pfc_device* dev = nullptr;
    hipStream_t stream = nullptr;
    struct Slot {
        pfc_job* job;
        int stage;          // 0 = needs a segment, 1 = walking, 2 = done
        uint64_t seg_hi;    // end of the current segment
        uint64_t span;      // current segment width
        uint32_t jcur;      // next multiplier step of the current segment
        uint32_t ncand;
    };
    std::vector<Slot> slots;
    pfc_batch_job* d_jobs = nullptr;
    size_t d_jobs_cap = 0;
    pfc_batch_range* d_ranges = nullptr;
    size_t d_ranges_cap = 0;
    uint32_t* d_words = nullptr;      // the round's device words (pfc::launch_batch), kBatchWords
    pfc_ctrl* h_ctrl = nullptr;       // pinned, one per slot
    size_t h_ctrl_cap = 0;
    uint64_t rounds = 0;
    // Captured device phase of a round (tuning.use_graph): uploads from the
    // pinned staging below, queue-head clear, the set-up, walk and fold
    // kernels, control-block gather and one readback, replayed every round;
    // captured again only when a buffer capacity or the launch shape changes.
    pfc_batch_job* h_jobs = nullptr;  // pinned staging, d_jobs_cap entries
    pfc_batch_range* h_ranges = nullptr;  // pinned staging, d_ranges_cap entries
    uint32_t* h_words = nullptr;      // pinned: d_words[6 .. 13) (the phases' range counts, the slot count)
    pfc_ctrl* d_ctrl_all = nullptr;   // gathered control blocks, d_jobs_cap entries
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    size_t graph_jobs_cap = 0, graph_ranges_cap = 0;
    const pfc_ctrl* graph_h_ctrl = nullptr;   // the readback target the graph was captured with
    uint32_t graph_k = 0;
    uint64_t graph_replays = 0;
};

namespace {

pfc_batch_job batch_view(pfc_job* j) {
    pfc_batch_job b{};
    b.modes = j->desc.modes;
    b.nmin = j->desc.nmin;
    b.nmax = j->desc.nmax;
    b.cap = j->plan.cap;
    b.P = j->d_p;
    b.V = j->d_v;
    b.ONE = j->d_one;
    b.RA = j->d_ra;
    b.RB = j->d_rb;
    b.ctrl = j->d_ctrl;
    b.fac = j->d_fac;
    b.fcap = (uint32_t)j->fcap;
    b.forms_off = j->tuning.walk_forms_off;
    b.ta = pfc_setup_table{j->d_ca, j->d_ga, j->ta.npos, j->ta.digit_bits};
    b.tb = pfc_setup_table{j->d_cb, j->d_gb, j->tb.npos, j->tb.digit_bits};
    b.list = j->d_list;
    b.composite = j->d_comp;
    b.clist = j->d_clist;
    b.cprefix = j->d_cprefix;
    return b;
}

template <class T>
int ensure_device(T** p, size_t* cap, size_t need) {
    if (need <= *cap) return PFC_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    size_t c = std::max<size_t>(need, 2 * *cap);
    hipError_t e = hipMalloc((void**)p, c * sizeof(T));
    if (e != hipSuccess) {
        *cap = 0;
        return hip_fail(e, "batcher allocation");
    }
    *cap = c;
    return PFC_OK;
}

// A pinned host array of at least `need` entries (contents not kept).
template <class T>
int ensure_pinned(T** p, size_t* cap, size_t need) {
    if (need <= *cap) return PFC_OK;
    if (*p) (void)hipHostFree(*p);
    *p = nullptr;
    const size_t c = std::max<size_t>(need, 2 * *cap);
    hipError_t e = hipHostMalloc((void**)p, c * sizeof(T), hipHostMallocDefault);
    if (e != hipSuccess) {
        *cap = 0;
        return hip_fail(e, "batcher pinned allocation");
    }
    *cap = c;
    return PFC_OK;
}

// Drains a finished segment into the job (factor records, sums, position).
// *redo: the segment overflowed the factor buffer and must be walked again
// in halves (nothing of it is kept).
int finish_segment(pfc_job* j, const pfc_ctrl& c, uint64_t seg_hi, uint64_t span, bool* redo) {
    *redo = false;
    int rc = check_tables(j);
    if (rc) return rc;
    if (c.flags & 2u) return set_error(PFC_ERR_VALIDATION, kReciprocalFailure);
    if (c.nfactor > j->fcap) {
        if (span > 1) {
            *redo = true;
            return PFC_OK;
        }
        char msg[128];
        snprintf(msg, sizeof(msg), "Error: number of results (%u) overflowed array.", c.nfactor);
        return set_error(PFC_ERR_FACTOR_OVERFLOW, msg);
    }
    rc = check_walked(j, c);
    if (rc) return rc;
    if (c.nfactor) {
        const size_t old = j->factors.size();
        j->factors.resize(old + c.nfactor);
        hipError_t e = hipMemcpy(j->factors.data() + old, j->d_fac, (size_t)c.nfactor * sizeof(pfc_factor),
                                 hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e, "factor copy");
    }
    j->dev_checksum += c.checksum;
    j->primecount += c.primecount;
    j->segments += 1;
    j->pos = seg_hi;
    return PFC_OK;
}

}  // namespace

extern "C" int pfc_batcher_create(pfc_device* dev, pfc_batcher** out) {
    pfc::g_err[0] = 0;
    if (!dev || !out) return set_error(PFC_ERR_INVALID, "batcher_create: null pointer");
    {
        const hipError_t se = hipSetDevice(dev->ordinal);
        if (se != hipSuccess) return hip_fail(se, "batcher_create: hipSetDevice");
    }
    pfc_batcher* b = new pfc_batcher();
    b->dev = dev;
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_words, pfc::kBatchWords * sizeof(uint32_t));
    if (e == hipSuccess) e = hipHostMalloc((void**)&b->h_words, pfc::kBatchWords * sizeof(uint32_t), hipHostMallocDefault);
    if (e != hipSuccess) {
        int rc = hip_fail(e, "batcher create");
        if (b->d_words) (void)hipFree(b->d_words);
        if (b->stream) (void)hipStreamDestroy(b->stream);
        delete b;
        return rc;
    }
    *out = b;
    return PFC_OK;
}

extern "C" int pfc_batcher_add(pfc_batcher* b, pfc_job* j) {
    if (!b || !j) return set_error(PFC_ERR_INVALID, "batcher_add: null pointer");
    if (j->dev != b->dev) return set_error(PFC_ERR_INVALID, "batcher_add: job on another device");
    pfc_batcher::Slot s{j, j->pos >= j->desc.pmax ? 2 : 0, 0, j->plan.span, 0, 0};
    b->slots.push_back(s);
    return PFC_OK;
}

extern "C" int pfc_batcher_pending(const pfc_batcher* b, size_t* n) {
    if (!b || !n) return set_error(PFC_ERR_INVALID, "batcher_pending: null pointer");
    *n = b->slots.size();
    return PFC_OK;
}

// Steps per segment of a work-budgeted round, at least this many.
constexpr uint32_t kMinRoundSteps = 256;

// The device phase of a round from one captured graph: the job views and
// ranges are staged in pinned host memory, which the graph's copy nodes read
// at each replay; the queue head is cleared by a kernel (no memset node);
// k_batch reads the range count from device memory and runs on a fixed grid
// (wavefronts past the queue's end return at once); every slot's control
// block is gathered into h_ctrl with one readback.  Captured again only when
// a buffer grows or the launch shape changes.
static int batcher_graph_round(pfc_batcher* b, const std::vector<pfc_batch_job>& jobs,
                               const std::vector<pfc_batch_range>& flat, const uint32_t nphase[6], uint32_t K) {
    hipStream_t st = b->stream;
    const size_t ns = jobs.size();
    size_t hj_cap = b->graph_jobs_cap, hr_cap = b->graph_ranges_cap;
    int rc = PFC_OK;
    const bool recapture = !b->gexec || b->graph_jobs_cap != b->d_jobs_cap ||
                           b->graph_ranges_cap != b->d_ranges_cap || b->graph_k != K ||
                           b->h_ctrl_cap < b->d_jobs_cap || b->graph_h_ctrl != b->h_ctrl;
    if (recapture) {
        if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->gexec = nullptr;
        b->graph = nullptr;
        // staging and gather buffers at the device buffers' capacities
        if (b->h_jobs) (void)hipHostFree(b->h_jobs);
        if (b->h_ranges) (void)hipHostFree(b->h_ranges);
        if (b->d_ctrl_all) (void)hipFree(b->d_ctrl_all);
        b->h_jobs = nullptr;
        b->h_ranges = nullptr;
        b->d_ctrl_all = nullptr;
        hj_cap = hr_cap = 0;
        size_t cc = 0;
        rc = ensure_pinned(&b->h_jobs, &hj_cap, b->d_jobs_cap);
        if (!rc) rc = ensure_pinned(&b->h_ranges, &hr_cap, b->d_ranges_cap);
        if (!rc) rc = ensure_device(&b->d_ctrl_all, &cc, b->d_jobs_cap);
        if (!rc && b->h_ctrl_cap < b->d_jobs_cap) {
            if (b->h_ctrl) (void)hipHostFree(b->h_ctrl);
            b->h_ctrl = nullptr;
            b->h_ctrl_cap = 0;
            rc = ensure_pinned(&b->h_ctrl, &b->h_ctrl_cap, b->d_jobs_cap);
        }
        if (rc) return rc;
        hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return hip_fail(e, "batcher hipStreamBeginCapture");
        e = hipMemcpyAsync(b->d_jobs, b->h_jobs, b->d_jobs_cap * sizeof(pfc_batch_job), hipMemcpyHostToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(b->d_ranges, b->h_ranges, b->d_ranges_cap * sizeof(pfc_batch_range),
                               hipMemcpyHostToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(b->d_words + 6, b->h_words, 7 * sizeof(uint32_t), hipMemcpyHostToDevice, st);
        if (e != hipSuccess) rc = hip_fail(e, "batcher graph upload");
        if (!rc) rc = pfc::launch_zero_heads(b->d_words, st);
        if (!rc) {
            // a fixed grid per phase: the kernels read their range counts
            // from d_words and return at once when a phase has none
            const uint32_t nb = (uint32_t)b->dev->prop.multiProcessorCount * 32;
            const uint32_t nblocks[6] = {nb, nb, nb, nb, nb, nb};
            rc = pfc::launch_batch(b->d_jobs, b->d_ranges, b->d_words, K, nblocks, st);
        }
        if (!rc) rc = pfc::launch_gather_ctrl(b->d_jobs, b->d_words + 12, b->d_ctrl_all, (uint32_t)b->d_jobs_cap, st);
        if (!rc) {
            e = hipMemcpyAsync(b->h_ctrl, b->d_ctrl_all, b->d_jobs_cap * sizeof(pfc_ctrl), hipMemcpyDeviceToHost, st);
            if (e != hipSuccess) rc = hip_fail(e, "batcher graph readback");
        }
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(e, "batcher hipStreamEndCapture");
        b->graph = g;
        e = hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) return hip_fail(e, "batcher hipGraphInstantiate");
        b->graph_jobs_cap = b->d_jobs_cap;
        b->graph_ranges_cap = b->d_ranges_cap;
        b->graph_h_ctrl = b->h_ctrl;
        b->graph_k = K;
    }
    memcpy(b->h_jobs, jobs.data(), ns * sizeof(pfc_batch_job));
    memcpy(b->h_ranges, flat.data(), flat.size() * sizeof(pfc_batch_range));
    for (int ph = 0; ph < 6; ++ph) b->h_words[ph] = nphase[ph];
    b->h_words[6] = (uint32_t)ns;
    hipError_t e = hipGraphLaunch(b->gexec, st);
    if (e != hipSuccess) return hip_fail(e, "batcher hipGraphLaunch");
    b->graph_replays += 1;
    return PFC_OK;
}

// One round; max_updates != 0 replaces max_steps by per-segment step counts
// that give about max_updates p*n updates over all walking segments (F|C
// counted twice), an equal share of them per segment (pfc::share_steps).
static int batcher_round(pfc_batcher* b, uint32_t max_steps, uint64_t max_updates, pfc_job** done,
                         size_t cap, size_t* ndone) {
    pfc::g_err[0] = 0;
    if (!b || !ndone) return set_error(PFC_ERR_INVALID, "batcher_round: null pointer");
    *ndone = 0;
    {
        const hipError_t se = hipSetDevice(b->dev->ordinal);
        if (se != hipSuccess) return hip_fail(se, "batcher_round: hipSetDevice");
    }
    hipStream_t st = b->stream;
    if (max_steps == 0) max_steps = 0xFFFFFFFFu;
    const size_t ns = b->slots.size();
    int rc = PFC_OK;
    if (b->h_ctrl_cap < ns) {
        if (b->h_ctrl) (void)hipHostFree(b->h_ctrl);
        b->h_ctrl = nullptr;
        b->h_ctrl_cap = std::max<size_t>(ns, 2 * b->h_ctrl_cap);
        hipError_t e = hipHostMalloc((void**)&b->h_ctrl, b->h_ctrl_cap * sizeof(pfc_ctrl), hipHostMallocDefault);
        if (e != hipSuccess) return hip_fail(e, "batcher pinned ctrl");
    }
    // 1. new segments
    bool any_gen = false;
    for (size_t i = 0; i < ns; ++i) {
        pfc_batcher::Slot& s = b->slots[i];
        if (s.stage != 0) continue;
        pfc_job* j = s.job;
        uint64_t hi = j->pos + s.span;
        if (hi > j->desc.pmax || hi < j->pos) hi = j->desc.pmax;
        s.seg_hi = hi;
        hipError_t e = hipMemsetAsync(j->d_ctrl, 0, sizeof(pfc_ctrl), st);
        if (e != hipSuccess) return hip_fail(e, "batcher ctrl clear");
        rc = pfc::launch_generate(j->pos, hi, j->d_p, j->plan.cap, j->d_ctrl, st, j->tuning.lds_segment);
        if (rc) return rc;
        e = hipMemcpyAsync(&b->h_ctrl[i], j->d_ctrl, sizeof(pfc_ctrl), hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return hip_fail(e, "batcher ctrl read");
        any_gen = true;
    }
    if (any_gen) {
        hipError_t e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(e, "batcher generate");
        for (size_t i = 0; i < ns; ++i) {
            pfc_batcher::Slot& s = b->slots[i];
            if (s.stage != 0) continue;
            const pfc_ctrl c = b->h_ctrl[i];
            if (c.flags & 1u) {  // more candidates than the bound: halve and retry next round
                if (s.span <= 1) return set_error(PFC_ERR_CAPACITY, "candidate buffer overflow");
                s.span /= 2;
                continue;
            }
            s.ncand = c.ncand;
            s.jcur = s.job->jbegin;
            s.stage = 1;
        }
    }
    uint32_t K, forced_blocks;
    {
        std::lock_guard<std::mutex> g(b->dev->mu);
        K = b->dev->tuning.primes_per_lane;
        forced_blocks = b->dev->tuning.batch_walk_blocks;
    }
    const uint64_t waves = (uint64_t)b->dev->prop.multiProcessorCount * 32;
    std::vector<uint32_t> slot_steps(ns, max_steps);
    if (max_updates) {
        std::vector<uint64_t> cost, cap, steps;
        std::vector<size_t> at;
        uint64_t total = 0;
        for (size_t i = 0; i < ns; ++i) {
            const pfc_batcher::Slot& s = b->slots[i];
            if (s.stage != 1) continue;
            at.push_back(i);
            cost.push_back((uint64_t)s.ncand * (s.job->d_rb ? 2 : 1));
            total += cost.back();
        }
        // A segment walks at most what one wavefront walks in such a round
        // with every wavefront busy (so that no item outlasts the round), or
        // the step count all segments would take if all took the same, if
        // that is more.
        const uint64_t wave_steps = std::max<uint64_t>(
            std::max<uint64_t>(kMinRoundSteps, max_updates / (waves * 64 * K)), total ? max_updates / total : 0);
        for (size_t k = 0; k < at.size(); ++k) {
            const pfc_batcher::Slot& s = b->slots[at[k]];
            cap.push_back(std::min<uint64_t>(s.job->jend - s.jcur, wave_steps));
        }
        steps.resize(at.size());
        pfc::share_steps(cost.data(), cap.data(), at.size(), max_updates, kMinRoundSteps, steps.data());
        for (size_t k = 0; k < at.size(); ++k)
            slot_steps[at[k]] = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(steps[k], 0xFFFFFFFFull));
    }
    // 2. work items: one range of items per segment (blocks of 64*K
    // candidates each), ranges ordered by their items' cost, longest first,
    // so that the queue ends on short items.
    std::vector<pfc_batch_job> jobs(ns);
    struct Tmp { pfc_batch_range r; uint32_t count; double cost, setup_cost; uint32_t ncand; };
    std::vector<Tmp> ranges;
    std::vector<uint8_t> finishing(ns, 0);
    uint64_t nitems = 0;
    // Walk items grow to several blocks only while the round keeps at least
    // four items per launched wavefront (the queue's tail stays short).
    uint64_t walk_blocks = 0;
    for (const pfc_batcher::Slot& s : b->slots)
        if (s.stage == 1) walk_blocks += (s.ncand + 64ull * K - 1) / (64ull * K);
    const uint32_t blocks_cap =
        (uint32_t)std::min<uint64_t>(pfc::kWalkMaxBlocks, std::max<uint64_t>(1, walk_blocks / (4 * waves)));
    for (size_t i = 0; i < ns; ++i) {
        pfc_batcher::Slot& s = b->slots[i];
        jobs[i] = batch_view(s.job);
        if (s.stage != 1) continue;
        pfc_job* j = s.job;
        const uint32_t j0 = s.jcur;
        const uint32_t j1 = (uint32_t)std::min<uint64_t>(j->jend, (uint64_t)j0 + slot_steps[i]);
        const uint32_t jw = launch_end(j, j0, j1);          // == j1 unless the test hook is set
        uint32_t flags = (j0 == j->jbegin ? 1u : 0u) | (j1 == j->jend ? 2u : 0u);
        if (s.ncand == 0) {
            finishing[i] = 1;  // nothing to walk; the segment is complete
            s.jcur = j->jend;
            continue;
        }
        // A walk item covers enough blocks of 64*K candidates that steps x
        // blocks reaches kWalkItemSteps: short rounds keep their items big.
        const uint32_t steps = std::max<uint32_t>(1u, jw - j0);
        const uint32_t blocks =
            forced_blocks ? forced_blocks : std::min<uint32_t>(blocks_cap, (pfc::kWalkItemSteps + steps - 1) / steps);
        const uint64_t per_item = 64ull * K * blocks;
        const uint32_t count = (uint32_t)((s.ncand + per_item - 1) / per_item);
        if (nitems + count > kMaxBatchItems) continue;   // left for the next round
        nitems += count;
        // Relative cost of one walk item: walk steps (the F|C walk takes ~1.9
        // steps per n) times its blocks.  The set-up list orders by
        // start-table size.
        const double per = (double)(j1 - j0) * ((j->d_rb) ? 1.9 : 1.0) * blocks;
        ranges.push_back({{(uint32_t)i, j0, jw, flags, walk_bound(j, jw), 0, blocks, 0}, count, per,
                          (double)(j->ta.chunks.size() + j->tb.chunks.size()), s.ncand});
        s.jcur = j1;
        if (j1 == j->jend) finishing[i] = 1;
    }
    std::stable_sort(ranges.begin(), ranges.end(), [](const Tmp& a, const Tmp& c) { return a.cost > c.cost; });
    bool gathered = false;                 // every slot's control block is in h_ctrl after the sync
    if (!ranges.empty()) {
        // Six lists in one array (pfc::launch_batch's phase order), each with
        // its own item ends: the walk ranges of walk mode 0..3 (longest items
        // first), the ranges of segments set up in this round (largest tables
        // first), the ranges of segments folded in it.
        std::vector<pfc_batch_range> flat;
        flat.reserve(3 * ranges.size());
        uint32_t nphase[6] = {0, 0, 0, 0, 0, 0};
        uint32_t nitem[6] = {0, 0, 0, 0, 0, 0};
        auto append = [&](const Tmp& t, int ph) {
            nitem[ph] += t.count;
            flat.push_back(t.r);
            flat.back().end = nitem[ph];
            nphase[ph] += 1;
        };
        for (int mode = 0; mode < 4; ++mode)
            for (const Tmp& t : ranges)
                if (pfc::walk_mode_of(b->slots[t.r.job].job->desc.modes) == mode) append(t, mode);
        std::vector<const Tmp*> setups;
        for (const Tmp& t : ranges)
            if (t.r.flags & 1u) setups.push_back(&t);
        std::stable_sort(setups.begin(), setups.end(),
                         [](const Tmp* a, const Tmp* c) { return a->setup_cost > c->setup_cost; });
        // set-up items are one block of 64*K candidates, fold items kFoldBlocks
        auto resized = [&](const Tmp& t, uint32_t blocks) {
            Tmp f = t;
            f.r.blocks = blocks;
            const uint64_t per_item = 64ull * K * blocks;
            f.count = (uint32_t)((t.ncand + per_item - 1) / per_item);
            return f;
        };
        for (const Tmp* t : setups) append(resized(*t, 1), 4);
        for (const Tmp& t : ranges)
            if (t.r.flags & 2u) append(resized(t, pfc::kFoldBlocks), 5);
        rc = ensure_device(&b->d_jobs, &b->d_jobs_cap, ns);
        if (!rc) rc = ensure_device(&b->d_ranges, &b->d_ranges_cap, flat.size());
        if (rc) return rc;
        bool use_graph;
        {
            std::lock_guard<std::mutex> g(b->dev->mu);
            use_graph = b->dev->tuning.use_graph != 0;
        }
        if (use_graph) {
            rc = batcher_graph_round(b, jobs, flat, nphase, K);
            if (rc) return rc;
            gathered = true;
        } else {
            for (int ph = 0; ph < 6; ++ph) b->h_words[ph] = nphase[ph];
            b->h_words[6] = (uint32_t)ns;
            hipError_t e = hipMemcpyAsync(b->d_jobs, jobs.data(), ns * sizeof(pfc_batch_job), hipMemcpyHostToDevice, st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(b->d_ranges, flat.data(), flat.size() * sizeof(pfc_batch_range), hipMemcpyHostToDevice,
                                   st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(b->d_words + 6, b->h_words, 7 * sizeof(uint32_t), hipMemcpyHostToDevice, st);
            if (e != hipSuccess) return hip_fail(e, "batcher upload");
            rc = pfc::launch_zero_heads(b->d_words, st);
            if (rc) return rc;
            const uint64_t nb = (uint64_t)b->dev->prop.multiProcessorCount * 32;
            uint32_t nblocks[6];
            for (int ph = 0; ph < 6; ++ph) nblocks[ph] = (uint32_t)std::min<uint64_t>(nitem[ph], nb);
            rc = pfc::launch_batch(b->d_jobs, b->d_ranges, b->d_words, K, nblocks, st);
            if (rc) return rc;
        }
    }
    // 3. drain finished segments
    for (size_t i = 0; i < ns && !gathered; ++i) {
        if (!finishing[i]) continue;
        hipError_t e = hipMemcpyAsync(&b->h_ctrl[i], b->slots[i].job->d_ctrl, sizeof(pfc_ctrl),
                                      hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return hip_fail(e, "batcher ctrl read");
    }
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "batcher round");
    b->rounds += 1;
    // Finished segments are drained; jobs with nothing left (including jobs
    // added with their position already at pmax) are handed back.
    std::vector<pfc_batcher::Slot> keep;
    for (size_t i = 0; i < ns; ++i) {
        pfc_batcher::Slot s = b->slots[i];
        if (finishing[i]) {
            bool redo = false;
            rc = finish_segment(s.job, b->h_ctrl[i], s.seg_hi, s.span, &redo);
            if (rc) return rc;
            if (redo) s.span /= 2;
            s.stage = s.job->pos >= s.job->desc.pmax ? 2 : 0;
        }
        if (s.stage == 2) {
            if (done && *ndone < cap) done[*ndone] = s.job;
            ++*ndone;
        } else {
            keep.push_back(s);
        }
    }
    b->slots.swap(keep);
    return PFC_OK;
}

extern "C" int pfc_batcher_round(pfc_batcher* b, uint32_t max_steps, pfc_job** done, size_t cap,
                                 size_t* ndone) {
    return batcher_round(b, max_steps, 0, done, cap, ndone);
}

extern "C" int pfc_batcher_round_work(pfc_batcher* b, uint64_t max_updates, pfc_job** done, size_t cap,
                                      size_t* ndone) {
    if (max_updates == 0) return set_error(PFC_ERR_INVALID, "batcher_round_work: max_updates is 0");
    return batcher_round(b, 0xFFFFFFFFu, max_updates, done, cap, ndone);
}

extern "C" int pfc_batcher_destroy(pfc_batcher* b) {
    if (!b) return PFC_OK;
    (void)hipSetDevice(b->dev->ordinal);
    (void)hipStreamSynchronize(b->stream);
    if (b->d_jobs) (void)hipFree(b->d_jobs);
    if (b->d_ranges) (void)hipFree(b->d_ranges);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->d_words) (void)hipFree(b->d_words);
    if (b->d_ctrl_all) (void)hipFree(b->d_ctrl_all);
    if (b->h_ctrl) (void)hipHostFree(b->h_ctrl);
    if (b->h_jobs) (void)hipHostFree(b->h_jobs);
    if (b->h_ranges) (void)hipHostFree(b->h_ranges);
    if (b->h_words) (void)hipHostFree(b->h_words);
    (void)hipStreamDestroy(b->stream);
    delete b;
    return PFC_OK;
}

extern "C" int pfc_batch_run(pfc_job** jobs, size_t njobs) {
    pfc::g_err[0] = 0;
    if (!jobs && njobs) return set_error(PFC_ERR_INVALID, "batch_run: null jobs");
    if (njobs == 0) return PFC_OK;
    pfc_batcher* b = nullptr;
    int rc = pfc_batcher_create(jobs[0]->dev, &b);
    if (rc) return rc;
    for (size_t i = 0; i < njobs && !rc; ++i) rc = pfc_batcher_add(b, jobs[i]);
    size_t left = njobs;
    while (!rc && left) {
        size_t nd = 0;
        rc = pfc_batcher_round(b, 0, nullptr, 0, &nd);
        left -= std::min(left, nd);
    }
    pfc_batcher_destroy(b);
    return rc;
}

// ====================================================================== harvest

extern "C" int pfc_is_prime(uint64_t p) {
    // Bases of verifyprime.cpp:111 (deterministic below 2^64).
    static const uint32_t bases[7] = {2, 325, 9375, 28178, 450775, 9780504, 1795265022};
    if (p < 2 || !(p & 1)) return p == 2;
    if (p < 9) return p != 1;
    const uint64_t pinv = h_inv(p);
    const uint64_t one = (uint64_t)(((u128)1 << 64) % p);
    const uint64_t r2 = (uint64_t)(((u128)one * one) % p);
    const uint64_t pm1 = p - one;  // Montgomery form of -1
    const int t = __builtin_ctzll(p - 1);
    const uint64_t dd = (p - 1) >> t;
    for (int i = 0; i < 7; ++i) {
        uint64_t b = bases[i] % p;
        if (b == 0) continue;
        uint64_t base = h_redc((u128)b * r2, p, pinv);
        uint64_t a = one;
        for (int k = 63 - __builtin_clzll(dd); k >= 0; --k) {
            a = h_redc((u128)a * a, p, pinv);
            if ((dd >> k) & 1) a = h_redc((u128)a * base, p, pinv);
        }
        if (a == one || a == pm1) continue;
        bool ok = false;
        for (int s = 1; s < t && !ok; ++s) {
            a = h_redc((u128)a * a, p, pinv);
            ok = a == pm1;
        }
        if (!ok) return 0;
    }
    return 1;
}

// X(n) mod p by direct products of the integers, primes or composites up to
// n, stopping once the product is 0 (then it can no longer be +-1).
static uint64_t slow_residue(uint64_t p, uint32_t n, int32_t type, const std::vector<uint8_t>& isp) {
    uint64_t r = 1 % p;
    for (uint32_t i = 2; i <= n && r; ++i) {
        if (type == PFC_PRIMORIAL && !isp[i]) continue;
        if (type == PFC_COMPOSITORIAL && isp[i]) continue;
        r = p < 0xFFFFFFFFull ? (r * i) % p : (uint64_t)(((u128)r * i) % p);
    }
    return r;
}

extern "C" int pfc_verify_factor(uint64_t p, uint32_t n, int32_t c, int32_t type) {
    std::vector<uint8_t> isp(n + 1, 0);
    if (type != PFC_FACTORIAL)
        for (uint32_t q : primes_in(2, n)) isp[q] = 1;
    const uint64_t r = slow_residue(p, n, type, isp);
    return (r == 1 && c == -1) || (r == p - 1 && c == 1);
}

extern "C" int pfc_harvest(pfc_factor* f, size_t n, int verify_threads, uint64_t* checksum,
                           size_t* nkept) {
    pfc::g_err[0] = 0;
    if ((!f && n) || !checksum || !nkept) return set_error(PFC_ERR_INVALID, "harvest: null pointer");
    std::sort(f, f + n, factor_less);
    if (verify_threads > 0 && n) {
        uint32_t maxn = 0;
        bool need_primes = false;
        for (size_t i = 0; i < n; ++i) {
            const uint32_t a = (uint32_t)(f[i].nc < 0 ? -f[i].nc : f[i].nc);
            maxn = std::max(maxn, a);
            need_primes |= f[i].type != PFC_FACTORIAL;
        }
        std::vector<uint8_t> isp(need_primes ? maxn + 1 : 1, 0);
        if (need_primes)
            for (uint32_t q : primes_in(2, maxn)) isp[q] = 1;
        std::atomic<size_t> next(0);
        std::atomic<long long> bad(-1);
        auto worker = [&]() {
            for (;;) {
                const size_t i = next.fetch_add(64);
                if (i >= n || bad.load() >= 0) return;
                for (size_t k = i; k < std::min(n, i + 64); ++k) {
                    const uint32_t a = (uint32_t)(f[k].nc < 0 ? -f[k].nc : f[k].nc);
                    const int32_t c = f[k].nc < 0 ? -1 : 1;
                    const uint64_t r = slow_residue(f[k].p, a, f[k].type, isp);
                    if (!((r == 1 && c == -1) || (r == f[k].p - 1 && c == 1))) {
                        long long exp = -1;
                        bad.compare_exchange_strong(exp, (long long)k);
                        return;
                    }
                }
            }
        };
        const int nt = std::max(1, std::min(verify_threads, 128));
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t) th.emplace_back(worker);
        for (auto& x : th) x.join();
        if (bad.load() >= 0) {
            const pfc_factor& b = f[bad.load()];
            const uint32_t a = (uint32_t)(b.nc < 0 ? -b.nc : b.nc);
            const char* form = b.type == PFC_FACTORIAL ? "!" : b.type == PFC_PRIMORIAL ? "#" : "!/#";
            char msg[200];
            snprintf(msg, sizeof(msg), "CPU factor verification failed!  %llu is not a factor of %u%s%+d",
                     (unsigned long long)b.p, a, form, b.nc < 0 ? -1 : 1);
            return set_error(PFC_ERR_VERIFY, msg);
        }
    }
    uint64_t last = 0;
    size_t k = 0;
    for (size_t i = 0; i < n; ++i) {
        if (f[i].p == last || pfc_is_prime(f[i].p)) {  // cl_sieve.cpp:561
            last = f[i].p;
            const uint32_t a = (uint32_t)(f[i].nc < 0 ? -f[i].nc : f[i].nc);
            const uint32_t c = f[i].nc < 0 ? 0xFFFFFFFFu : 1u;
            *checksum += (uint64_t)(uint32_t)(a + c);  // cl_sieve.cpp:583
            f[k++] = f[i];
        }
    }
    *nkept = k;
    return PFC_OK;
}

extern "C" int pfc_format_factor(const pfc_factor* f, char* buf, size_t cap) {
    if (!f || !buf) return set_error(PFC_ERR_INVALID, "format: null pointer");
    const uint32_t a = (uint32_t)(f->nc < 0 ? -f->nc : f->nc);
    const int c = f->nc < 0 ? -1 : 1;
    const char* form = f->type == PFC_FACTORIAL ? "!" : f->type == PFC_PRIMORIAL ? "#" : "!/#";
    return snprintf(buf, cap, "%llu | %u%s%+d\n", (unsigned long long)f->p, a, form, c);
}

// putil.c:34-116 grammar: strtoull base 0, then one suffix: K M G T P (x1000
// each), eN (x10^N), k m g t p (x2^10 each), bN (x2^N).  0 ok, -1 parse
// error, -2 out of [lo, hi].
extern "C" int pfc_parse_uint64(const char* s, uint64_t lo, uint64_t hi, uint64_t* out) {
    if (!s || !out) return -1;
    char* tail = nullptr;
    errno = 0;
    uint64_t v = strtoull(s, &tail, 0);
    if (errno != 0 || v > hi) return -2;
    const char c = *tail;
    static const char dec[] = "KMGTP", bin[] = "kmgtp";
    const char* pd = c ? strchr(dec, c) : nullptr;
    const char* pb = c ? strchr(bin, c) : nullptr;
    if (pd) {
        if (tail[1] != '\0') return -1;
        for (int k = 0; k <= (int)(pd - dec); ++k) {
            if (v > hi / 1000) return -2;
            v *= 1000;
        }
    } else if (pb) {
        if (tail[1] != '\0') return -1;
        const int sh = 10 * (int)(pb - bin + 1);
        if (v > (hi >> sh)) return -2;
        v <<= sh;
    } else if (c == 'e' || c == 'E' || c == 'b' || c == 'B') {
        errno = 0;
        char* t2 = nullptr;
        unsigned long ex = strtoul(tail + 1, &t2, 0);
        if (errno != 0) return -2;
        if (*t2 != '\0') return -1;
        const bool ten = (c == 'e' || c == 'E');
        while (ex-- > 0) {
            if (ten) {
                if (v > hi / 10) return -2;
                v *= 10;
            } else {
                if (v > (hi >> 1)) return -2;
                v <<= 1;
            }
        }
    } else if (c != '\0') {
        return -1;
    }
    if (v < lo) return -2;
    *out = v;
    return 0;
}

// ====================================================================== tables (C ABI)

extern "C" int pfc_build_setup_table(uint32_t kind, uint32_t N, uint64_t* chunks, size_t chunk_cap,
                                     size_t* nchunks, uint32_t* groups, size_t group_cap, size_t* ngroups,
                                     uint32_t* npos, uint32_t* digit_bits) {
    if (kind > 2 || !nchunks || !ngroups || !npos || !digit_bits)
        return set_error(PFC_ERR_INVALID, "build_setup_table: bad args");
    HostTable T = build_table(kind, N);
    *nchunks = T.chunks.size();
    *ngroups = T.groups.size();
    *npos = T.npos;
    *digit_bits = T.digit_bits;
    if (chunks) memcpy(chunks, T.chunks.data(), std::min(chunk_cap, T.chunks.size()) * 8);
    if (groups) memcpy(groups, T.groups.data(), std::min(group_cap, T.groups.size()) * 4);
    return PFC_OK;
}
