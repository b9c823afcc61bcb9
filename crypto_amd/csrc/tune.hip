// Autotuner of the MSM and NTT schedules, and the text file that caches its
// tables.  Every candidate runs on synthetic device inputs and must reproduce
// the built-in schedule's output byte for byte before its time counts.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "engine.h"
#include "kinet/crypto/hip_runtime.h"
#include "kinet/gpukit/gpukit.h"
#include "msm_internal.h"
#include "ntt.h"
#include "scan.h"

namespace kz {

void fixed_base_mul(int curve, const uint32_t* d_scalars, size_t n, uint8_t* d_out_wire, hipStream_t s);

namespace {

struct Mismatch : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Seeded 256-bit values below the scalar field modulus of `curve` (top byte
// masked: BLS12-381 r < 2^255 starts 0x73, BN254 r starts 0x30; secp256k1's
// group order is within 2^129 of 2^256, so its scalars take all 256 bits),
// uploaded.
void fill_scalars(int curve, uint64_t seed, size_t n, void* d, hipStream_t s) {
    std::vector<uint64_t> h(n * 4);
    uint64_t x = seed * 0x9E3779B97F4A7C15ull + 1;
    for (auto& w : h) {
        x += 0x9E3779B97F4A7C15ull;  // splitmix64
        uint64_t z = x;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        w = z ^ (z >> 31);
    }
    const uint64_t top = curve == 0 ? ~0ull : curve == 2 ? 0x3FFFFFFFFFFFFFFFull : 0x1FFFFFFFFFFFFFFFull;
    for (size_t i = 0; i < n; ++i) h[4 * i + 3] &= top;
    KZ_HIP(hipMemcpyAsync(d, h.data(), n * 32, hipMemcpyHostToDevice, s));
    KZ_HIP(hipStreamSynchronize(s));  // h is freed on return
}

__global__ void k_words_differ(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b, size_t n,
                               int* __restrict__ flag) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        if (a[i] != b[i]) {
            atomicOr(flag, 1);
            return;
        }
}

bool device_equal(const void* a, const void* b, size_t bytes, int* d_flag, hipStream_t s) {
    device_fill_u32(reinterpret_cast<uint32_t*>(d_flag), 0, 1, s);
    const size_t words = bytes / 4;
    const unsigned grid = (unsigned)std::min<size_t>(ceil_div(words, 256), 4096);
    hipLaunchKernelGGL(k_words_differ, dim3(grid ? grid : 1), dim3(256), 0, s, static_cast<const uint32_t*>(a),
                       static_cast<const uint32_t*>(b), words, d_flag);
    KZ_HIP(hipGetLastError());
    int h = 0;
    KZ_HIP(hipMemcpyAsync(&h, d_flag, sizeof h, hipMemcpyDeviceToHost, s));
    KZ_HIP(hipStreamSynchronize(s));
    return h == 0;
}

// Mean milliseconds of fn() on stream s: one untimed warm-up, then five
// trials of at least two calls and about 20 ms each; the fastest trial counts
// (a single call of a ~7 ms MSM varies by several per cent from run to run).
template <class Fn>
float time_ms(Fn&& fn, hipStream_t s) {
    hipEvent_t a, b;
    KZ_HIP(hipEventCreate(&a));
    KZ_HIP(hipEventCreate(&b));
    float best = 0;
    try {
        KZ_HIP(hipEventRecord(a, s));
        fn();
        KZ_HIP(hipEventRecord(b, s));
        KZ_HIP(hipEventSynchronize(b));
        float first = 0;
        KZ_HIP(hipEventElapsedTime(&first, a, b));
        int reps = first > 0 ? (int)(20.0f / first) + 1 : 50;
        if (reps < 2) reps = 2;
        if (reps > 50) reps = 50;
        for (int t = 0; t < 5; ++t) {
            KZ_HIP(hipEventRecord(a, s));
            for (int r = 0; r < reps; ++r) fn();
            KZ_HIP(hipEventRecord(b, s));
            KZ_HIP(hipEventSynchronize(b));
            float ms = 0;
            KZ_HIP(hipEventElapsedTime(&ms, a, b));
            ms /= (float)reps;
            if (t == 0 || ms < best) best = ms;
        }
    } catch (...) {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        throw;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return best;
}

int floor_log2(size_t n) {
    int lg = 0;
    while (((size_t)1 << (lg + 1)) <= n) ++lg;
    return lg;
}

// Synthetic MSM inputs: n multiples of the generator (so in G1) and enough
// scalars for `nsc` points.
struct MsmInputs {
    DevBuf wire, pts, inf, sc, ksc;
};

void make_msm_inputs(int curve, size_t n, size_t nsc, MsmInputs& in, hipStream_t s) {
    in.ksc.ensure(n * 32);
    fill_scalars(curve, 11, n, in.ksc.ptr, s);
    in.wire.ensure(n * (curve == 2 ? 96 : 64));
    fixed_base_mul(curve, in.ksc.as<uint32_t>(), n, in.wire.as<uint8_t>(), s);
    in.pts.ensure(n * (size_t)affine_words(curve) * 4);
    in.inf.ensure(n);
    points_from_wire(curve, in.wire.as<uint8_t>(), in.pts.as<uint32_t>(), in.inf.as<uint8_t>(), n, s);
    in.sc.ensure(nsc * 32);
    fill_scalars(curve, 12, nsc, in.sc.ptr, s);
    KZ_HIP(hipStreamSynchronize(s));
}

struct Choice {
    int c = 0, seg = 0;
    float ms = 0;
};

// One MSM of 2^lg points (modes plain and GLV): window bits, then segments.
Choice tune_single(Engine& e, int curve, int mode, int lg, const MsmInputs& in, DevBuf& out) {
    const size_t n = (size_t)1 << lg;
    const size_t ob = msm_out_bytes(curve);
    hipStream_t s = e.stream;
    out.ensure(2 * ob);
    uint8_t* d_ref = out.as<uint8_t>();
    uint8_t* d_out = d_ref + ob;
    auto run = [&](int c, int seg, uint8_t* dst, bool glv = true) {
        MsmParams prm;
        prm.c = c;
        prm.seg = seg;
        prm.g1 = glv && mode == kMsmGlv;
        msm_device(curve, in.pts.as<uint32_t>(), in.inf.as<uint8_t>(), in.sc.as<uint32_t>(), n, prm, e.ws, s, dst);
    };
    // built-in schedule: the reference output
    int oc = 0, oseg = 0;
    msm_get_tuned(curve, mode, lg, oc, oseg);
    msm_set_tuned(curve, mode, lg, 0, 0);
    const int c_rule = msm_window_bits(curve, n, mode);
    msm_set_tuned(curve, mode, lg, oc, oseg);
    run(c_rule, 0, d_ref);
    std::vector<uint8_t> ref(ob), got(ob);
    KZ_HIP(hipMemcpyAsync(ref.data(), d_ref, ob, hipMemcpyDeviceToHost, s));
    KZ_HIP(hipStreamSynchronize(s));
    auto check = [&](int c, int seg) {
        KZ_HIP(hipMemcpyAsync(got.data(), d_out, ob, hipMemcpyDeviceToHost, s));
        KZ_HIP(hipStreamSynchronize(s));
        if (got != ref)
            throw Mismatch("MSM schedule c=" + std::to_string(c) + " seg=" + std::to_string(seg) +
                           " differs from the built-in schedule at 2^" + std::to_string(lg));
    };
    const bool verbose = std::getenv("KINET_HIP_TUNE_VERBOSE") != nullptr;
    const size_t msm_n = mode == kMsmGlv ? 2 * n : n;
    const int lo = std::max(5, lg - 9), hi = std::min(mode == kMsmGlv ? 16 : 18, std::max(5, lg - 1 + (mode == kMsmGlv)));
    Choice best;
    for (int c = lo; c <= hi; ++c) {
        if (mode == kMsmGlv && !lds_sort_supported(c, msm_n)) continue;
        const float ms = time_ms([&] { run(c, 0, d_out); }, s);
        check(c, 0);
        if (verbose) std::fprintf(stderr, "msm curve=%d mode=%d 2^%d c=%d seg=rule %.4f ms\n", curve, mode, lg, c, ms);
        if (best.c == 0 || ms < best.ms) best = Choice{c, 0, ms};
    }
    if (best.c == 0) throw std::invalid_argument("no window width applies to this size class");
    for (int seg : {16, 8, 4, 2}) {
        if (seg > (1 << (best.c - 1))) continue;
        const float ms = time_ms([&] { run(best.c, seg, d_out); }, s);
        check(best.c, seg);
        if (verbose)
            std::fprintf(stderr, "msm curve=%d mode=%d 2^%d c=%d seg=%d %.4f ms\n", curve, mode, lg, best.c, seg, ms);
        if (ms < best.ms) {  // best.seg 0: the built-in rule, kept when fastest
            best.seg = seg;
            best.ms = ms;
        }
    }
    if (mode == kMsmGlv) {
        // the plain MSM under its own cell of this size class, kept when it beats the split
        const float ms = time_ms([&] { run(0, 0, d_out, false); }, s);
        check(0, 0);
        if (verbose) std::fprintf(stderr, "msm curve=%d mode=%d 2^%d plain %.4f ms\n", curve, mode, lg, ms);
        if (ms < best.ms) best = Choice{kMsmGlvOff, 0, ms};
    }
    return best;
}

// Batches of equal jobs of 2^lg points, >= 2^20 points in all (>= 4 jobs).
Choice tune_batched(Engine& e, int curve, int lg, const MsmInputs& in, size_t jobs, DevBuf& out, int* d_flag) {
    const size_t n = (size_t)1 << lg;
    const size_t ob = msm_out_bytes(curve);
    hipStream_t s = e.stream;
    out.ensure(2 * ob * jobs);
    uint8_t* d_ref = out.as<uint8_t>();
    uint8_t* d_out = d_ref + ob * jobs;
    MsmWorkspace ws;
    auto plan_of = [&](int c, MsmBatchPlan& p) {
        std::vector<MsmJob> specs(jobs);
        for (size_t j = 0; j < jobs; ++j) {
            specs[j] = MsmJob{};
            specs[j].n = (uint32_t)n;
            specs[j].sc_base = j * n;
            specs[j].pt_base = 0;
            specs[j].c = (uint32_t)c;
        }
        msm_batch_plan(curve, specs.data(), jobs, 0, p);
        msm_batch_reserve(p, ws);
    };
    int oc = 0, oseg = 0;
    msm_get_tuned(curve, kMsmBatched, lg, oc, oseg);
    msm_set_tuned(curve, kMsmBatched, lg, 0, 0);
    const int c_rule = msm_window_bits(curve, n, kMsmBatched);
    msm_set_tuned(curve, kMsmBatched, lg, oc, oseg);
    {
        MsmBatchPlan p;
        plan_of(c_rule, p);
        msm_batch_run(p, in.pts.as<uint32_t>(), in.inf.as<uint8_t>(), in.sc.as<uint32_t>(), ws, s, d_ref);
        KZ_HIP(hipStreamSynchronize(s));
    }
    Choice best;
    for (int c = std::max(5, lg - 9); c <= std::min(16, std::max(5, lg - 1)); ++c) {
        MsmBatchPlan p;
        plan_of(c, p);
        const float ms = time_ms(
            [&] {
                msm_batch_run(p, in.pts.as<uint32_t>(), in.inf.as<uint8_t>(), in.sc.as<uint32_t>(), ws, s, d_out);
            },
            s);
        if (!device_equal(d_ref, d_out, ob * jobs, d_flag, s))
            throw Mismatch("batched MSM window bits " + std::to_string(c) +
                           " differ from the built-in schedule at 2^" + std::to_string(lg));
        if (std::getenv("KINET_HIP_TUNE_VERBOSE"))
            std::fprintf(stderr, "msm curve=%d batched 2^%d x %zu jobs c=%d %.4f ms\n", curve, lg, jobs, c, ms);
        if (best.c == 0 || ms < best.ms) best = Choice{c, 0, ms};
    }
    KZ_HIP(hipStreamSynchronize(s));  // ws is freed on return
    return best;
}

void tune_msm(int curve, int mode, int log_min, int log_max) {
    if (!msm_tunable(curve, mode)) throw std::invalid_argument("curve or mode");
    if (log_min < 1 || log_max > 24 || log_min > log_max) throw std::invalid_argument("size classes out of range");
    if (mode == kMsmGlv && log_min < 11) throw std::invalid_argument("GLV size classes start at 2^11");
    Engine& e = engine();
    const size_t nmax = (size_t)1 << log_max;
    auto jobs_of = [](int lg) { return std::max<size_t>(4, ((size_t)1 << 20) >> lg); };
    size_t nsc = nmax;
    if (mode == kMsmBatched)
        for (int lg = log_min; lg <= log_max; ++lg) nsc = std::max(nsc, jobs_of(lg) << lg);
    MsmInputs in;
    DevBuf out, flag;
    flag.ensure(sizeof(int));
    make_msm_inputs(curve, nmax, nsc, in, e.stream);  // takes the engine lock itself
    std::lock_guard<std::mutex> g(e.mu);
    for (int lg = log_min; lg <= log_max; ++lg) {
        const Choice ch = mode == kMsmBatched ? tune_batched(e, curve, lg, in, jobs_of(lg), out, flag.as<int>())
                                              : tune_single(e, curve, mode, lg, in, out);
        msm_set_tuned(curve, mode, lg, ch.c, ch.seg);
    }
    KZ_HIP(hipStreamSynchronize(e.stream));
}

void tune_ntt(int field, int log_min, int log_max) {
    if (field != 0 && field != 1) throw std::invalid_argument("field");
    if (log_min < 1 || log_max > 24 || log_min > log_max) throw std::invalid_argument("size classes out of range");
    Engine& e = engine();
    std::lock_guard<std::mutex> g(e.mu);
    hipStream_t s = e.stream;
    const size_t total = (size_t)1 << std::max(18, log_max);
    DevBuf in, ref, out, tmp, flag;
    flag.ensure(sizeof(int));
    in.ensure(total * 32);
    ref.ensure(total * 32);
    out.ensure(total * 32);
    tmp.ensure(total * 32);
    fill_scalars(field == 0 ? 2 : 1, 13, total, in.ptr, s);  // BLS12-381 / BN254 Fr
    int* d_err = flag.as<int>();
    for (int lg = log_min; lg <= log_max; ++lg) {
        // a batch of 2^18 elements in all (at most 2^14 transforms, the grid's y
        // extent is 65535)
        const size_t n = (size_t)1 << lg, count = std::min<size_t>(16384, std::max<size_t>(1, ((size_t)1 << 18) >> lg));
        int o[3];
        ntt_get_schedule(field, lg, o[0], o[1], o[2]);
        ntt_set_schedule(field, lg, 0, 0, 0);
        ntt_prepare(field, lg, false, s);
        auto run = [&](uint32_t* dst) {
            ntt_device_batch(field, lg, false, in.as<uint32_t>(), dst, tmp.as<uint32_t>(), count, d_err, s);
        };
        run(ref.as<uint32_t>());
        // each trial's four-step tables are freed after it (up to n x 36 bytes
        // per split); the chosen split's are rebuilt below
        const size_t mark = ntt_table_mark();
        int bp[3] = {0, 0, 0};
        float best = 0;
        const int pm = kNttPassMaxLog;
        for (int p1 = 1; p1 <= std::min(pm, lg); ++p1)
            for (int p2 = 0; p2 <= std::min(pm, lg - p1); ++p2)
                for (int p3 = 0; p3 <= std::min(pm, lg - p1 - p2); ++p3) {
                    if (!ntt_split_valid(lg, p1, p2, p3)) continue;
                    ntt_set_schedule(field, lg, p1, p2, p3);
                    const float ms = time_ms([&] { run(out.as<uint32_t>()); }, s);
                    const bool same = device_equal(ref.ptr, out.ptr, count * n * 32, d_err, s);
                    KZ_HIP(hipStreamSynchronize(s));
                    ntt_drop_tables_since(mark);
                    if (!same) {
                        ntt_set_schedule(field, lg, o[0], o[1], o[2]);
                        throw Mismatch("NTT split " + std::to_string(p1) + "/" + std::to_string(p2) + "/" +
                                       std::to_string(p3) + " differs from the built-in split");
                    }
                    if (bp[0] == 0 || ms < best) {
                        best = ms;
                        bp[0] = p1;
                        bp[1] = p2;
                        bp[2] = p3;
                    }
                }
        ntt_set_schedule(field, lg, bp[0], bp[1], bp[2]);
        ntt_prepare(field, lg, false, s);
    }
    KZ_HIP(hipStreamSynchronize(s));
}

// ------------------------------------------------------------ table files
struct Cell {
    bool msm;
    int a, b, lg, x, y, z;  // msm: curve mode lg c seg; ntt: field lg p1 p2 p3
};

std::vector<Cell> parse_table(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open tuning file " + path);
    std::vector<Cell> cells;
    std::string line;
    int no = 0;
    while (std::getline(f, line)) {
        ++no;
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        std::istringstream ls(line);
        std::string kind;
        if (!(ls >> kind)) continue;
        Cell c{};
        bool ok = false;
        std::string extra;
        if (kind == "msm") {
            c.msm = true;
            ok = (bool)(ls >> c.a >> c.b >> c.lg >> c.x >> c.y) && !(ls >> extra);
            ok = ok && msm_tunable(c.a, c.b) && c.lg >= 0 && c.lg < 32 &&
                 (c.x == 0 || (c.x >= 5 && c.x <= 18) || (c.b == kMsmGlv && c.x == kMsmGlvOff)) &&
                 (c.y == 0 || (c.y > 0 && c.y <= 16 && (c.y & (c.y - 1)) == 0));
        } else if (kind == "ntt") {
            c.msm = false;
            ok = (bool)(ls >> c.a >> c.lg >> c.x >> c.y >> c.z) && !(ls >> extra);
            ok = ok && (c.a == 0 || c.a == 1) && c.lg >= 0 && c.lg <= kNttMaxLog &&
                 ((c.x == 0 && c.y == 0 && c.z == 0) || ntt_split_valid(c.lg, c.x, c.y, c.z));
        }
        if (!ok) throw std::invalid_argument(path + ":" + std::to_string(no) + ": malformed tuning line");
        cells.push_back(c);
    }
    return cells;
}

void load_table(const std::string& path) {
    for (const Cell& c : parse_table(path)) {
        if (c.msm) msm_set_tuned(c.a, c.b, c.lg, c.x, c.y);
        else ntt_set_schedule(c.a, c.lg, c.x, c.y, c.z);
    }
}

void save_table(const std::string& path) {
    std::ostringstream o;
    o << "# MSM and NTT schedules per size class k = floor(log2 n)\n"
      << "# msm <curve 0=secp256k1 1=BN254 2=BLS12-381> <mode 0 single 1 batched 2 GLV> <k> <window bits> "
         "<segment>\n"
      << "# (a GLV cell of window bits 1: the plain MSM is faster at that size class)\n"
      << "# ntt <field 0=BLS12-381 Fr 1=BN254 Fr> <k> <log2 pass 1> <pass 2> <pass 3>\n";
    for (int curve = 0; curve <= 2; ++curve)
        for (int mode = 0; mode < kMsmModes; ++mode)
            for (int lg = 0; lg < 32 && msm_tunable(curve, mode); ++lg) {
                int c = 0, seg = 0;
                msm_get_tuned(curve, mode, lg, c, seg);
                if (c || seg) o << "msm " << curve << ' ' << mode << ' ' << lg << ' ' << c << ' ' << seg << '\n';
            }
    for (int field = 0; field <= 1; ++field)
        for (int lg = 0; lg <= kNttMaxLog; ++lg) {
            int p[3];
            ntt_get_schedule(field, lg, p[0], p[1], p[2]);
            if (p[0] | p[1] | p[2])
                o << "ntt " << field << ' ' << lg << ' ' << p[0] << ' ' << p[1] << ' ' << p[2] << '\n';
        }
    std::ofstream f(path, std::ios::trunc);
    if (!f || !(f << o.str()) || !f.flush()) throw std::runtime_error("cannot write tuning file " + path);
}

void clear_tables() {
    for (int curve = 0; curve <= 2; ++curve)
        for (int mode = 0; mode < kMsmModes; ++mode)
            for (int lg = 0; lg < 32 && msm_tunable(curve, mode); ++lg) msm_set_tuned(curve, mode, lg, 0, 0);
    for (int field = 0; field <= 1; ++field)
        for (int lg = 0; lg <= kNttMaxLog; ++lg) ntt_set_schedule(field, lg, 0, 0, 0);
}

template <class Fn>
int guarded(Fn&& fn) {
    try {
        fn();
        return GPUKIT_OK;
    } catch (const Mismatch& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_DETERMINISM;
    } catch (const std::invalid_argument& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BAD_SIZE;
    } catch (const std::exception& ex) {
        set_last_error(ex.what());
        return GPUKIT_ERR_BACKEND;
    }
}

}  // namespace

void tune_load_env() {
    const char* f = std::getenv("KINET_HIP_TUNE_FILE");
    if (f && *f) load_table(f);
}

}  // namespace kz

using namespace kz;

extern "C" {

int kinet_hip_tune_msm(uint32_t curve, int mode, int log_min, int log_max) {
    return guarded([&] { tune_msm((int)curve, mode, log_min, log_max); });
}

int kinet_hip_tune_ntt(uint32_t field, int log_min, int log_max) {
    return guarded([&] { tune_ntt((int)field, log_min, log_max); });
}

int kinet_hip_tune_save(const char* path) {
    if (!path) return GPUKIT_ERR_NULL_ARG;
    return guarded([&] { save_table(path); });
}

int kinet_hip_tune_load(const char* path) {
    if (!path) return GPUKIT_ERR_NULL_ARG;
    return guarded([&] { load_table(path); });
}

void kinet_hip_tune_clear(void) { (void)guarded([] { clear_tables(); }); }

int kinet_hip_tune_get_msm(uint32_t curve, int mode, int log_n, int* c, int* seg) {
    if (!c || !seg) return GPUKIT_ERR_NULL_ARG;
    return guarded([&] { msm_get_tuned((int)curve, mode, log_n, *c, *seg); });
}

int kinet_hip_tune_get_ntt(uint32_t field, int log_n, int* p1, int* p2, int* p3) {
    if (!p1 || !p2 || !p3) return GPUKIT_ERR_NULL_ARG;
    return guarded([&] { ntt_schedule_of((int)field, log_n, *p1, *p2, *p3); });
}

}  // extern "C"
