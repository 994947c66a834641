// pfcsieve_main.cpp -- the `pfcsieve` command-line program over libpfcsieve.so.
//
// Keeps the reference's command-line surface (main.cpp:27-188): -! -# -c,
// -n/-N, -p/-P with the putil number grammar, -v verify threads, -s/--test
// self-test, --device (accepted, ignored); factors.txt lines and final
// checksum line (cl_sieve.cpp:565-577, 772-822); A/B checkpoint files with an
// additive state sum (cl_sieve.cpp:104-243).  BOINC is not linked: the program
// behaves as the reference does standalone, with the device ordinal taken
// from the slot's init_data.xml when one is present (main.cpp:252-274);
// trickle-up, fraction-done and critical sections have no client to talk to
// and are not emitted.
#include <getopt.h>
#include <inttypes.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <string>
#include <vector>

#include "../../include/pfc.h"

namespace {

const char* kResults = "factors.txt";
const char* kStateA = "stateA.ckp";
const char* kStateB = "stateB.ckp";

// Checkpoint record, field for field the reference's workStatus
// (cl_sieve.h:14-18), so state files are interchangeable on x86-64.
struct WorkStatus {
    uint64_t pmin, pmax, p, checksum, primecount, factorcount, last_trickle, state_sum;
    uint32_t nmin, nmax;
    bool factorial, primorial, compositorial;
};
static_assert(sizeof(WorkStatus) == 80, "workStatus layout");

struct Options {
    WorkStatus st{};
    uint32_t threads = 2;  // main.cpp:217
    bool test = false;
    int device = 0;
};

uint64_t state_sum(const WorkStatus& s) {
    return s.pmin + s.pmax + s.p + s.checksum + s.primecount + s.factorcount + s.last_trickle + s.nmin +
           s.nmax;
}

void say(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void say(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vprintf(fmt, ap);
    va_end(ap);
    fflush(stdout);
}

[[noreturn]] void die(const char* msg) {
    fprintf(stderr, "%s\n", msg);
    printf("%s\n", msg);
    exit(EXIT_FAILURE);
}

void help() {
    say("Welcome to PFCSieve (MI355X build), a program to find factors of primorial, factorial, and "
        "compositorial prime candidates of the form n#+-1, n!+-1, and n!/#+-1\n");
    say("Program usage:\n");
    say("-!\tUse factorial mode\n-#\tUse primorial mode\n-c\tUse compositorial mode\n");
    say("\t\tNote: -! -c can be used together to find factors of both at the same time.\n");
    say("-n #\tStart primorial n#+-1, factorial n!+-1, or compositorial n!/#+-1\n");
    say("-N #\tEnd primorial N#+-1, factorial N!+-1, or compositorial N!/#+-1\n");
    say("\t\tN range is 101 <= -n < -N < 2^31, [-n, -N) exclusive\n");
    say("-p #\tStarting prime factor p\n-P #\tEnd prime factor P\n");
    say("\t\tP range is 3 <= -p < -P < 2^64, [-p, -P) exclusive\n");
    say("-v #\tOptional, specify the number of CPU threads used to verify factors.  Default is 2, max is 128.\n");
    say("-s \tPerform self test to verify proper operation of the program with the current GPU.\n");
    say("-h\tPrint this help\n");
    exit(EXIT_FAILURE);
}

int parse_u64(const char* s, uint64_t lo, uint64_t hi, uint64_t* out) { return pfc_parse_uint64(s, lo, hi, out); }

int parse_u32(const char* s, uint32_t lo, uint32_t hi, uint32_t* out) {
    uint64_t v = 0;
    int rc = pfc_parse_uint64(s, lo, hi, &v);
    if (rc == 0) *out = (uint32_t)v;
    return rc;
}

Options parse_args(int argc, char** argv) {
    Options o;
    static const struct option longs[] = {{"device", optional_argument, 0, 'd'},
                                          {"test", no_argument, 0, 's'},
                                          {0, 0, 0, 0}};
    int opt, ind = -1;
    while ((opt = getopt_long(argc, argv, "p:P:n:N:v:s!#cfh", longs, &ind)) != -1) {
        int rc = 0;
        switch (opt) {  // main.cpp:52-112
            case 'p': rc = parse_u64(optarg, 3, UINT64_MAX - 1, &o.st.pmin); break;
            case 'P': rc = parse_u64(optarg, 4, UINT64_MAX, &o.st.pmax); break;
            case 'n': rc = parse_u32(optarg, 101, 0x7FFFFFFE, &o.st.nmin); break;
            case 'N': rc = parse_u32(optarg, 102, 0x7FFFFFFF, &o.st.nmax); break;
            case 'v': rc = parse_u32(optarg, 1, 128, &o.threads); break;
            case 's': o.test = true; say("Performing self test.\n"); break;
            case '!': o.st.factorial = true; say("\n-! argument specified for factorial mode.\n\n"); break;
            case '#': o.st.primorial = true; say("\n-# argument specified for primorial mode.\n\n"); break;
            case 'c': o.st.compositorial = true; say("\n-c argument specified for compositorial mode.\n\n"); break;
            case 'f': case 'd': break;
            default: help();
        }
        if (rc == -1) {
            say("%s: invalid argument -%c %s\n", argv[0], opt, optarg ? optarg : "");
            exit(EXIT_FAILURE);
        }
        if (rc == -2) {
            say("%s: out of range argument -%c %s\n", argv[0], opt, optarg ? optarg : "");
            exit(EXIT_FAILURE);
        }
        ind = -1;
    }
    return o;
}

uint32_t modes_of(const WorkStatus& s) {
    return (s.factorial ? PFC_MODE_FACTORIAL : 0) | (s.primorial ? PFC_MODE_PRIMORIAL : 0) |
           (s.compositorial ? PFC_MODE_COMPOSITORIAL : 0);
}

// ---------------------------------------------------------------- checkpoint files

bool read_one(const char* name, const WorkStatus& want, WorkStatus* out) {
    FILE* f = fopen(name, "rb");
    if (!f) return false;
    WorkStatus s;
    bool ok = fread(&s, sizeof(s), 1, f) == 1;
    fclose(f);
    if (!ok) {
        say("Cannot parse %s !!!\n", name);
        return false;
    }
    if (s.pmin != want.pmin || s.pmax != want.pmax || s.nmin != want.nmin || s.nmax != want.nmax ||
        s.factorial != want.factorial || s.primorial != want.primorial || s.compositorial != want.compositorial) {
        say("Invalid checkpoint file %s !!!\n", name);
        return false;
    }
    if (state_sum(s) != s.state_sum) {
        say("Checksum error in %s !!!\n", name);
        return false;
    }
    *out = s;
    return true;
}

// read_state (cl_sieve.cpp:139-233): keep the valid file with the larger p.
bool read_state(WorkStatus* st, bool* write_a_next) {
    WorkStatus a, b;
    bool ga = read_one(kStateA, *st, &a), gb = read_one(kStateB, *st, &b);
    if (ga && gb) {
        if (a.p > b.p) gb = false;
        else ga = false;
    }
    if (ga) { *st = a; *write_a_next = false; say("Resuming from checkpoint in %s\n", kStateA); return true; }
    if (gb) { *st = b; *write_a_next = true; say("Resuming from checkpoint in %s\n", kStateB); return true; }
    return false;
}

void write_state(WorkStatus* st, bool* write_a_next) {
    st->state_sum = state_sum(*st);
    const char* name = *write_a_next ? kStateA : kStateB;
    FILE* f = fopen(name, "wb");
    if (!f) { fprintf(stderr, "Cannot open %s !!!\n", name); return; }
    bool ok = fwrite(st, sizeof(*st), 1, f) == 1;
    if (fclose(f) == 0 && ok) *write_a_next = !*write_a_next;
    else fprintf(stderr, "Cannot write checkpoint to file. Continuing...\n");
}

// ---------------------------------------------------------------- one search

uint32_t seg_candidates_env = 0;

void check_rc(int rc) {
    if (rc != PFC_OK) die(pfc_last_error());
}

// Harvests the job's factor records: CPU verify, isPrime filter, factors.txt
// lines, checksum fold (getResults, cl_sieve.cpp:450-593).
void harvest(pfc_job* job, WorkStatus* st, uint32_t threads) {
    size_t n = 0;
    check_rc(pfc_job_factors(job, nullptr, 0, &n));
    if (!n) return;
    std::vector<pfc_factor> f(n);
    check_rc(pfc_job_factors(job, f.data(), n, &n));
    say("processing %zu factors on CPU\n", n);
    size_t kept = 0;
    uint64_t cs = st->checksum;
    check_rc(pfc_harvest(f.data(), n, (int)threads, &cs, &kept));
    say("Verified %zu factors.\n", n);
    FILE* out = fopen(kResults, "a");
    if (!out) die("Cannot open factors.txt !!!");
    char line[128];
    for (size_t i = 0; i < kept; ++i) {
        pfc_format_factor(&f[i], line, sizeof(line));
        if (fputs(line, out) < 0) die("Cannot write to factors.txt !!!");
    }
    fclose(out);
    if (kept < n) say("discarded %zu 2-PRP factor records\n", n - kept);
    st->checksum = cs;
    st->factorcount += kept;
    check_rc(pfc_job_clear_factors(job));
}

// finalizeResults (cl_sieve.cpp:772-822).
void finalize(const WorkStatus& st) {
    if (st.factorcount) {
        FILE* f = fopen(kResults, "r");
        if (!f) die("Cannot open factors.txt !!!");
        uint64_t lc = 0;
        char line[256];
        while (fgets(line, sizeof(line), f)) ++lc;
        fclose(f);
        if (lc < st.factorcount) die("ERROR: Missing factors in factors.txt !!!");
    }
    FILE* f = fopen(kResults, "a");
    if (!f) die("Cannot open factors.txt !!!");
    if (st.factorcount) fprintf(f, "%016" PRIX64 "\n", st.checksum);
    else fprintf(f, "no factors\n%016" PRIX64 "\n", st.checksum);
    fclose(f);
}

void sieve(pfc_device* dev, WorkStatus& st, uint32_t threads, bool test) {
    st.p = st.pmin;
    pfc_job_desc desc{};
    desc.modes = modes_of(st);
    desc.nmin = st.nmin;
    desc.nmax = st.nmax;
    desc.pmin = st.pmin;
    desc.pmax = st.pmax;
    if (st.factorial && st.compositorial && !st.primorial)
        say("Sieving for factors of factorial and compositorial\n");
    pfc_job* job = nullptr;
    check_rc(pfc_job_create(dev, &desc, &job));
    say("Starting sieve at p: %" PRIu64 " n: %u\nStopping sieve at P: %" PRIu64 " N: %u\n", st.pmin, st.nmin,
        st.pmax, st.nmax);
    bool write_a = true;
    if (!test && read_state(&st, &write_a)) {
        say("Resuming from checkpoint, current p: %" PRIu64 "\n", st.p);
        if (st.p == st.pmax) {
            say("Workunit complete.\n");
            pfc_job_destroy(job);
            exit(EXIT_SUCCESS);
        }
        check_rc(pfc_job_seek(job, st.p, st.checksum, st.primecount));
    } else {
        FILE* f = fopen(kResults, "w");
        if (!f) die("Cannot open factors.txt !!!");
        fclose(f);
        st.last_trickle = (uint64_t)time(nullptr);
        st.checksum = 0;
        st.primecount = 0;
        st.factorcount = 0;
        check_rc(pfc_job_seek(job, st.p, 0, 0));
    }
    // Environment knobs (not reference flags): checkpoint interval in seconds
    // (default 60, cl_sieve.cpp:1581) and, for resume tests, a stop after a
    // number of segments that leaves the last checkpoint behind.
    const char* e_ck = getenv("PFC_CHECKPOINT_SECONDS");
    const char* e_stop = getenv("PFC_STOP_AFTER_SEGMENTS");
    const long ck_every = e_ck ? atol(e_ck) : 60;
    const long stop_after = (e_stop && !test) ? atol(e_stop) : 0;
    long segs = 0;
    time_t t0 = time(nullptr), last_ckpt = t0, last_prog = t0;
    // The job's checksum accumulator carries the harvested n+c parts too.
    uint64_t factor_part = 0;
    for (;;) {
        uint64_t pos = 0;
        check_rc(pfc_job_position(job, &pos));
        if (pos >= st.pmax) break;
        check_rc(pfc_job_run(job, 1));
        time_t now = time(nullptr);
        if (now - last_prog > 1) {
            check_rc(pfc_job_position(job, &pos));
            say("Sieve Progress: %.1f%%\n", 100.0 * (double)(pos - st.pmin) / (double)(st.pmax - st.pmin));
            last_prog = now;
        }
        ++segs;
        if (now - last_ckpt >= ck_every || (stop_after && segs >= stop_after)) {  // checkpoint at a segment boundary
            pfc_job_stats js;
            check_rc(pfc_job_stats_get(job, &js));
            st.checksum = js.dev_checksum + factor_part;
            harvest(job, &st, threads);
            factor_part = st.checksum - js.dev_checksum;
            check_rc(pfc_job_position(job, &st.p));
            st.primecount = js.primecount;
            write_state(&st, &write_a);
            say("Checkpoint, current p: %" PRIu64 "\n", st.p);
            last_ckpt = now;
            if (stop_after && segs >= stop_after) {
                say("Stopping after %ld segments (PFC_STOP_AFTER_SEGMENTS).\n", segs);
                pfc_job_destroy(job);
                exit(EXIT_SUCCESS);
            }
        }
    }
    pfc_job_stats js;
    check_rc(pfc_job_stats_get(job, &js));
    st.checksum = js.dev_checksum + factor_part;
    harvest(job, &st, threads);
    st.p = st.pmax;
    st.primecount = js.primecount;
    say("Sieve Progress: %.1f%%\n", 100.0);
    write_state(&st, &write_a);
    finalize(st);
    pfc_job_destroy(job);
    say("Sieve finished in %d sec.\n", (int)(time(nullptr) - t0));
    say("factors %" PRIu64 ", prime count %" PRIu64 ", checksum %016" PRIX64 "\n", st.factorcount, st.primecount,
        st.checksum);
}

// The reference's 16 self-test ranges and golden results (cl_sieve.cpp:1787-2080).
struct TestCase {
    uint32_t modes;
    uint64_t pmin, pmax;
    uint32_t nmin, nmax;
    uint64_t factorcount, primecount, checksum;
};
const TestCase kTests[16] = {
    {1, 100000000ull, 101000000ull, 1000000, 2000000, 1071, 54211, 0x000004F844B5103Cull},
    {1, 1000000000000ull, 1000010000000ull, 10000, 2000000, 3, 361727, 0x0505A1C238896511ull},
    {1, 101ull, 100000ull, 101, 1000000, 42821, 9571, 0x0000000065DDB8A0ull},
    {1, 1000000000000ull, 1000001000000ull, 100000000, 110000000, 3, 36249, 0x00804FE7D7AA6C09ull},
    {2, 100000000ull, 101000000ull, 101, 25000000, 1703, 54211, 0x0000027EFF497990ull},
    {2, 101ull, 2000000ull, 101, 2000000, 24503, 148954, 0x000000027BF5B8E0ull},
    {2, 100000000000ull, 100005000000ull, 9000000, 110000000, 32, 197222, 0x0022FE7C09210B4Bull},
    {2, 1730720716000000ull, 1730720720000000ull, 600000, 30000000, 1, 114208, 0x5CDCB47F7E9532C2ull},
    {4, 200000000ull, 200010000ull, 101, 26000000, 127, 529, 0x0000001848D8AFBBull},
    {4, 101ull, 100000ull, 101, 1000000, 34271, 9571, 0x000000006FF88EAEull},
    {4, 200000000000ull, 200005000000ull, 15000000, 20000000, 13, 192386, 0x0088B59C23CD3E2Bull},
    {4, 1000000000000ull, 1000001000000ull, 700000, 25000000, 2, 36249, 0x0080997AF3BF42FEull},
    {5, 100000000000ull, 100010000000ull, 96000, 2000000, 27, 394403, 0x00D214CC0EF0ECB4ull},
    {5, 101ull, 110000ull, 101, 1000000, 84077, 10433, 0x00000000EFB634E9ull},
    {5, 10100000000ull, 10101000000ull, 11500000, 12500000, 19, 43374, 0x0002578EA9FD63C7ull},
    {5, 2000000000000ull, 2000020000000ull, 670000, 2460000, 3, 706162, 0x1D63BBC574E8D50Full},
};

int run_test(pfc_device* dev, uint32_t threads) {
    say("Beginning self test of 16 ranges.\n");
    time_t t0 = time(nullptr);
    int good = 0;
    for (int i = 0; i < 16; ++i) {
        const TestCase& c = kTests[i];
        WorkStatus st{};
        st.factorial = c.modes & 1;
        st.primorial = c.modes & 2;
        st.compositorial = c.modes & 4;
        st.pmin = c.pmin;
        st.pmax = c.pmax;
        st.nmin = c.nmin;
        st.nmax = c.nmax;
        sieve(dev, st, threads, true);
        const bool ok = st.factorcount == c.factorcount && st.primecount == c.primecount && st.checksum == c.checksum;
        say("test case %d %s.\n\n", i + 1, ok ? "passed" : "failed");
        good += ok;
    }
    if (good == 16) say("All test cases completed successfully!\n");
    else say("Self test FAILED!\n");
    say("Elapsed time: %d sec.\n", (int)(time(nullptr) - t0));
    return good == 16 ? EXIT_SUCCESS : EXIT_FAILURE;
}

// Device ordinal as boinc_get_opencl_ids finds it (main.cpp:252-274): the
// <gpu_device_num> (else <gpu_opencl_dev_index>) of init_data.xml in the
// working directory; without the file, device 0.
int device_from_init_data() {
    FILE* f = fopen("init_data.xml", "rb");
    if (!f) {
        say("init_data.xml not found, using device 0.\n");
        return 0;
    }
    std::string xml;
    char buf[4096];
    size_t k;
    while ((k = fread(buf, 1, sizeof(buf), f)) > 0) xml.append(buf, k);
    fclose(f);
    for (const char* tag : {"<gpu_device_num>", "<gpu_opencl_dev_index>"}) {
        const size_t at = xml.find(tag);
        if (at == std::string::npos) continue;
        const long v = strtol(xml.c_str() + at + strlen(tag), nullptr, 10);
        if (v >= 0 && v < 1024) {
            say("init_data.xml: using device %ld.\n", v);
            return (int)v;
        }
    }
    say("init_data.xml has no GPU device number, using device 0.\n");
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    say("\nPFCSieve (MI355X gfx950 build)\n");
    Options o = parse_args(argc, argv);
    if (!o.test && !o.st.factorial && !o.st.primorial && !o.st.compositorial)
        die("\n-! or -# or -c argument is required\nuse -h for help");
    if (const char* sc = getenv("PFC_SEG_CANDIDATES")) seg_candidates_env = (uint32_t)atol(sc);
    o.device = device_from_init_data();
    pfc_device* dev = nullptr;
    check_rc(pfc_device_open(o.device, &dev));
    if (seg_candidates_env) {
        pfc_tuning t{};
        t.seg_candidates = seg_candidates_env;
        check_rc(pfc_set_tuning(dev, &t));
    }
    char name[128];
    int cus = 0;
    uint64_t hbm = 0;
    pfc_device_info(dev, name, sizeof(name), &cus, &hbm);
    say("GPU Info:\n  Name: \t\t%s\n  Compute Units: \t%d\n  HBM: \t\t\t%" PRIu64 " MiB\n", name, cus, hbm >> 20);
    int rc = EXIT_SUCCESS;
    if (o.test) {
        rc = run_test(dev, o.threads);
    } else {
        sieve(dev, o.st, o.threads, false);
    }
    pfc_device_close(dev);
    return rc;
}
