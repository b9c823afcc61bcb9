// Host build of the library's device arithmetic (crypto_amd/csrc/field.h,
// curve.h), exported for the CPU tests, which compare it with the oracle.
// Test infrastructure only.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <map>
#include <stdexcept>
#include <cstring>
#include <thread>
#include <vector>

#include "curve.h"
#include "curve_ed.h"
#include "curve_quad.h"
#include "field_row.h"
#include "glv.h"
#include "ntt_math.h"
#include "serve_core.h"

using namespace kz;

namespace {

template <class F>
Fe<F> from_le(const uint8_t* b, int nbytes) {
    uint32_t w[16] = {0};
    for (int k = 0; k < nbytes; ++k) w[k / 4] |= (uint32_t)b[k] << (8 * (k % 4));
    return fe_from_words<F>(w, (nbytes + 3) / 4);
}

template <class F>
void to_le(const Fe<F>& a, uint8_t* b, int nbytes) {
    uint32_t w[16] = {0};
    fe_to_words<F>(a, w, (nbytes + 3) / 4);
    for (int k = 0; k < nbytes; ++k) b[k] = (uint8_t)(w[k / 4] >> (8 * (k % 4)));
}

// op: 0 add 1 sub 2 mul 3 inv 4 square of a; canonical LE in/out
template <class F>
void field_op(int op, const uint8_t* a, const uint8_t* b, uint8_t* out, int nbytes) {
    Fe<F> x = fe_to_mont<F>(from_le<F>(a, nbytes)), y = fe_to_mont<F>(from_le<F>(b, nbytes)), r;
    switch (op) {
        case 0: r = fe_add<F>(x, y); break;
        case 1: r = fe_sub<F>(x, y); break;
        case 2: r = fe_mul<F>(x, y); break;
        case 4: r = fe_sqr<F>(x); break;
        default: r = fe_inv<F>(x); break;
    }
    to_le<F>(fe_from_mont<F>(r), out, nbytes);
}

template <class C>
Xyzz<typename C::Fp> read_wire(const uint8_t* w) {
    using F = typename C::Fp;
    bool any = false;
    for (int i = 0; i < 2 * C::FB; ++i) any |= w[i] != 0;
    if (!any) return xyzz_inf<F>();
    Affine<F> a;
    a.x = fe_to_mont<F>(fe_from_be_bytes<F>(w, C::FB));
    a.y = fe_to_mont<F>(fe_from_be_bytes<F>(w + C::FB, C::FB));
    return xyzz_from_affine<F>(a);
}

template <class C>
void write_wire(const Xyzz<typename C::Fp>& p, uint8_t* w) {
    using F = typename C::Fp;
    Affine<F> a;
    bool inf;
    xyzz_to_affine<F>(p, a, inf);
    if (inf) { std::memset(w, 0, 2 * C::FB); return; }
    fe_to_be_bytes<F>(fe_from_mont<F>(a.x), w, C::FB);
    fe_to_be_bytes<F>(fe_from_mont<F>(a.y), w + C::FB, C::FB);
}

// op: 0 general add, 1 mixed add (b affine), 2 double a
template <class C>
void point_op(int op, const uint8_t* a, const uint8_t* b, uint8_t* out) {
    using F = typename C::Fp;
    Xyzz<F> p = read_wire<C>(a), r;
    if (op == 0) {
        r = xyzz_add<F>(p, read_wire<C>(b));
    } else if (op == 1) {
        Xyzz<F> q = read_wire<C>(b);
        Affine<F> qa;
        qa.x = q.X;
        qa.y = q.Y;
        r = p;
        xyzz_madd<F>(r, qa);
    } else if (op == 3) {  // p + 3 q through three mixed additions of the accumulation loop
        Xyzz<F> q = read_wire<C>(b);
        Affine<F> qa;
        qa.x = q.X;
        qa.y = q.Y;
        r = p;
        for (int k = 0; k < 3; ++k) xyzz_madd_lazy<F>(r, qa);
        r = xyzz_canon_acc<F>(r);
    } else if (op == 5) {  // p + 20 q: the loop's bounds in their steady state
        Xyzz<F> q = read_wire<C>(b);
        Affine<F> qa;
        qa.x = q.X;
        qa.y = q.Y;
        r = p;
        for (int k = 0; k < 20; ++k) xyzz_madd_lazy<F>(r, qa);
        r = xyzz_canon_acc<F>(r);
    } else {
        r = xyzz_dbl<F>(p);
    }
    write_wire<C>(r, out);
}

// ---------------------------------------------------------------- quads
// Four host threads in lock step stand in for the four lanes of a quad: the
// device's quad-cooperative point operations (curve_quad.h) run unchanged
// with this lane exchange in place of ds_swizzle.
struct SpinBarrier {
    std::atomic<int> count{0}, gen{0};
    void wait() {
        const int g = gen.load(std::memory_order_acquire);
        if (count.fetch_add(1, std::memory_order_acq_rel) == 3) {
            count.store(0, std::memory_order_relaxed);
            gen.fetch_add(1, std::memory_order_acq_rel);
        } else {
            for (int spins = 0; gen.load(std::memory_order_acquire) == g; ++spins)
                if (spins > 4096) std::this_thread::yield();
        }
    }
};

SpinBarrier g_quad_bar;
uint32_t g_quad_slot[4][16];

struct HostQuad {
    static thread_local int id;
    static int lane() { return id; }
    template <int K, class F>
    static Fe<F> read(const Fe<F>& a) {
        std::memcpy(g_quad_slot[id], a.v, sizeof a.v);
        g_quad_bar.wait();
        Fe<F> r;
        std::memcpy(r.v, g_quad_slot[K], sizeof r.v);
        g_quad_bar.wait();
        return r;
    }
};
thread_local int HostQuad::id = 0;

// Runs fn(lane) on four lanes; true if their results (returned as Xyzz)
// agree, with the result in out.
template <class F, class Fn, class Pt = Xyzz<F>>
bool run_quad(Fn fn, Pt& out) {
    Pt res[4];
    std::thread t[3];
    for (int k = 1; k < 4; ++k)
        t[k - 1] = std::thread([&, k] {
            HostQuad::id = k;
            res[k] = fn();
        });
    HostQuad::id = 0;
    res[0] = fn();
    for (auto& th : t) th.join();
    for (int k = 1; k < 4; ++k)
        if (std::memcmp(&res[k], &res[0], sizeof res[0]) != 0) return false;
    out = res[0];
    return true;
}

// value + k p without reduction (value < p, k small: stays below 2^(29 N))
template <class F>
Fe<F> add_kp(const Fe<F>& a, int k) {
    uint32_t w[16] = {0}, pw[16] = {0};
    const int nw = (kLimbBits * F::N + 31) / 32;
    fe_to_words<F>(a, w, nw);
    fe_to_words<F>(fe_const<F>(F::P), pw, nw);
    for (int t = 0; t < k; ++t) {
        uint64_t c = 0;
        for (int i = 0; i < nw; ++i) {
            c += (uint64_t)w[i] + pw[i];
            w[i] = (uint32_t)c;
            c >>= 32;
        }
    }
    return fe_from_words<F>(w, nw);
}

// XYZZ form of an affine wire point with ZZ = z^2, ZZZ = z^3 (z LE bytes),
// each coordinate raised by k[i] p (the lazy forms the device passes along).
template <class C>
Xyzz<typename C::Fp> lazy_point(const uint8_t* wire, const uint8_t* z_le, const int* k) {
    using F = typename C::Fp;
    Xyzz<F> p = read_wire<C>(wire);
    if (xyzz_is_inf<F>(p)) return p;
    const Fe<F> z = fe_to_mont<F>(from_le<F>(z_le, C::FB));
    const Fe<F> z2 = fe_mul<F>(z, z), z3 = fe_mul<F>(z2, z);
    Xyzz<F> r;
    r.X = add_kp<F>(fe_mul<F>(p.X, z2), k[0]);
    r.Y = add_kp<F>(fe_mul<F>(p.Y, z3), k[1]);
    r.ZZ = add_kp<F>(z2, k[2]);
    r.ZZZ = add_kp<F>(z3, k[3]);
    return r;
}

template <class C>
int quad_op(int op, int lazy, const uint8_t* a, const uint8_t* b, const uint8_t* za, const uint8_t* zb, const int* ka,
            const int* kb, uint8_t* out) {
    using F = typename C::Fp;
    const Xyzz<F> P = lazy_point<C>(a, za, ka), Q = lazy_point<C>(b, zb, kb);
    Xyzz<F> r;
    bool same;
    if (lazy)
        same = run_quad<F>([&] { return op == 0 ? xyzz_add_quad<F, MulCallLazy, HostQuad>(P, Q)
                                                : xyzz_dbl_quad<F, MulCallLazy, HostQuad>(P); }, r);
    else
        same = run_quad<F>([&] { return op == 0 ? xyzz_add_quad<F, MulCall, HostQuad>(P, Q)
                                                : xyzz_dbl_quad<F, MulCall, HostQuad>(P); }, r);
    write_wire<C>(xyzz_canon_acc<F>(r), out);
    return same ? 0 : -1;
}

// Sum of n points (n a power of two) by levels of lazy quad additions with no
// canonical step in between: at each level element i takes element i + half,
// the shape of a fold tree (three levels for fan-in 8); then canonical.
template <class C>
int quad_tree(int lazy, const uint8_t* pts, const uint8_t* zs, int n, uint8_t* out) {
    using F = typename C::Fp;
    std::vector<Xyzz<F>> v(n);
    const int k0[4] = {0, 0, 0, 0};
    for (int i = 0; i < n; ++i) v[i] = lazy_point<C>(pts + 2 * C::FB * i, zs + C::FB * i, k0);
    bool same = true;
    for (int half = n / 2; half >= 1; half /= 2)
        for (int i = 0; i < half; ++i) {
            Xyzz<F> r;
            const Xyzz<F> A = v[i], B = v[i + half];
            if (lazy)
                same &= run_quad<F>([&] { return xyzz_add_quad<F, MulCallLazy, HostQuad>(A, B); }, r);
            else
                same &= run_quad<F>([&] { return xyzz_add_quad<F, MulCall, HostQuad>(A, B); }, r);
            v[i] = r;
        }
    write_wire<C>(xyzz_canon_acc<F>(v[0]), out);
    return same ? 0 : -1;
}

// count lazy quad doublings of a point (the window combine's chain), then canonical
template <class C>
int quad_dbl_chain(const uint8_t* a, const uint8_t* za, const int* ka, int count, uint8_t* out) {
    using F = typename C::Fp;
    const Xyzz<F> P = lazy_point<C>(a, za, ka);
    Xyzz<F> r;
    const bool same = run_quad<F>([&] {
        Xyzz<F> x = P;
        for (int i = 0; i < count; ++i) x = xyzz_dbl_quad<F, MulCallLazy, HostQuad>(x);
        return x;
    }, r);
    write_wire<C>(xyzz_canon_acc<F>(r), out);
    return same ? 0 : -1;
}

// ---------------------------------------------------------------- rows
// A 64-lane wave in host arrays with the lane operations of WaveDev
// (field_row.h): the row-distributed product and point operations run
// unchanged.  DPP row semantics as WaveDev documents them.
struct WaveHost {
    struct V {
        uint32_t l[64];
    };
    struct V64 {
        uint64_t l[64];
    };
    struct B {
        bool l[64];
    };
    template <class T, class Fn>
    static T map(Fn fn) {
        T r;
        for (int i = 0; i < 64; ++i) r.l[i] = fn(i);
        return r;
    }
    static V lane16() { return map<V>([](int i) { return (uint32_t)(i & 15); }); }
    static V row() { return map<V>([](int i) { return (uint32_t)((i >> 4) & 3); }); }
    static V c(uint32_t x) { return map<V>([&](int) { return x; }); }
    static V add(V a, V b) { return map<V>([&](int i) { return a.l[i] + b.l[i]; }); }
    static V sub(V a, V b) { return map<V>([&](int i) { return a.l[i] - b.l[i]; }); }
    static V and_(V a, uint32_t m) { return map<V>([&](int i) { return a.l[i] & m; }); }
    static V shl(V a, int s) { return map<V>([&](int i) { return a.l[i] << s; }); }
    static V mulc(V a, uint32_t k) { return map<V>([&](int i) { return a.l[i] * k; }); }
    static V sra(V a, int s) { return map<V>([&](int i) { return (uint32_t)((int32_t)a.l[i] >> s); }); }
    static V64 wide(V a) { return map<V64>([&](int i) { return (uint64_t)a.l[i]; }); }
    static V lo(V64 a) { return map<V>([&](int i) { return (uint32_t)a.l[i]; }); }
    static V hi(V64 a) { return map<V>([&](int i) { return (uint32_t)(a.l[i] >> 32); }); }
    static V64 mk64(V l, V h) { return map<V64>([&](int i) { return ((uint64_t)h.l[i] << 32) | l.l[i]; }); }
    static V64 add64(V64 a, V64 b) { return map<V64>([&](int i) { return a.l[i] + b.l[i]; }); }
    static V64 shr64(V64 a, int s) { return map<V64>([&](int i) { return a.l[i] >> s; }); }
    static V64 mad(V a, V b, V64 c) { return map<V64>([&](int i) { return (uint64_t)a.l[i] * b.l[i] + c.l[i]; }); }
    static V64 madc(V a, uint32_t b, V64 c) { return map<V64>([&](int i) { return (uint64_t)a.l[i] * b + c.l[i]; }); }
    static B eqc(V a, uint32_t k) { return map<B>([&](int i) { return a.l[i] == k; }); }
    static B ltc(V a, uint32_t k) { return map<B>([&](int i) { return a.l[i] < k; }); }
    static B gtc(V a, uint32_t k) { return map<B>([&](int i) { return a.l[i] > k; }); }
    static B ne(V a, V b) { return map<B>([&](int i) { return a.l[i] != b.l[i]; }); }
    static B or_(B a, B b) { return map<B>([&](int i) { return a.l[i] || b.l[i]; }); }
    static B ge(V a, V b) { return map<B>([&](int i) { return a.l[i] >= b.l[i]; }); }
    static B gt(V a, V b) { return map<B>([&](int i) { return a.l[i] > b.l[i]; }); }
    static V shr32(V a, int s) { return map<V>([&](int i) { return a.l[i] >> s; }); }
    static uint64_t ballot(B b) {
        uint64_t m = 0;
        for (int i = 0; i < 64; ++i) m |= (uint64_t)b.l[i] << i;
        return m;
    }
    static V rowbits(uint64_t m) { return map<V>([&](int i) { return (uint32_t)(m >> (i & 48)) & 0xFFFFu; }); }
    static V sel(B b, V x, V y) { return map<V>([&](int i) { return b.l[i] ? x.l[i] : y.l[i]; }); }
    static V64 sel64(B b, V64 x, V64 y) { return map<V64>([&](int i) { return b.l[i] ? x.l[i] : y.l[i]; }); }
    template <int K>
    static V bcast(V x) {
        return map<V>([&](int i) { return x.l[(i & ~15) | K]; });
    }
    template <int S>
    static V ror(V x) {
        return map<V>([&](int i) { return x.l[(i & ~15) | (((i & 15) - S) & 15)]; });
    }
    template <int S>
    static V shr(V x) {
        return map<V>([&](int i) { return (i & 15) < S ? 0u : x.l[i - S]; });
    }
    template <int K>
    static V rowread(V x) {
        return map<V>([&](int i) { return x.l[(K << 4) | (i & 15)]; });
    }
    static bool any(B b) {
        for (int i = 0; i < 64; ++i)
            if (b.l[i]) return true;
        return false;
    }
    static uint32_t readlane(V x, int l) { return x.l[l]; }
};

// Every row holds the same limbs, and lanes N..15 hold zero.
template <class F>
bool row_consistent(const FeRow<WaveHost, F>& a) {
    for (int i = 0; i < 64; ++i) {
        if ((i & 15) >= F::N && a.v.l[i]) return false;
        if (a.v.l[i] != a.v.l[i & 15]) return false;
    }
    return true;
}

// Four products at once (row q: a[16 q ..] times b[16 q ..], N limbs each),
// and fe_mul_k<F, 1, false> of the same pairs.  -1 if a row's lanes N..15
// are not zero.
template <class F>
int row_mul_check(const uint32_t* a, const uint32_t* b, uint32_t* out_row, uint32_t* out_ref) {
    using W = WaveHost;
    const RowConsts<W, F> K = row_consts<W, F>();
    FeRow<W, F> x, y;
    for (int i = 0; i < 64; ++i) {
        const bool in = (i & 15) < F::N;
        x.v.l[i] = in ? a[i] : 0;
        y.v.l[i] = in ? b[i] : 0;
    }
    const FeRow<W, F> r = row_mul<W, F>(x, y, K);
    int bad = 0;
    for (int i = 0; i < 64; ++i) {
        out_row[i] = r.v.l[i];
        if ((i & 15) >= F::N && r.v.l[i]) bad = 1;
    }
    for (int q = 0; q < 4; ++q) {
        Fe<F> fa, fb;
        for (int j = 0; j < F::N; ++j) {
            fa.v[j] = a[16 * q + j];
            fb.v[j] = b[16 * q + j];
        }
        const Fe<F> m = fe_mul_k<F, 1, false>(fa, fb);
        for (int j = 0; j < 16; ++j) out_ref[16 * q + j] = j < F::N ? m.v[j] : 0;
    }
    return bad ? -1 : 0;
}

template <class F>
void raw_xyzz(const Xyzz<F>& p, uint32_t* out) {
    for (int j = 0; j < F::N; ++j) {
        out[j] = p.X.v[j];
        out[F::N + j] = p.Y.v[j];
        out[2 * F::N + j] = p.ZZ.v[j];
        out[3 * F::N + j] = p.ZZZ.v[j];
    }
}

// op 0: a + b, 1: 2a, by the row operations and by the lazy quad ones; raw
// coordinates of both (4 N limbs each) and the canonical wire result of the
// rows.  -1 if the rows disagree.
template <class C>
int row_op(int op, const uint8_t* a, const uint8_t* b, const uint8_t* za, const uint8_t* zb, const int* ka,
           const int* kb, uint8_t* out, uint32_t* raw_row, uint32_t* raw_quad) {
    using F = typename C::Fp;
    using W = WaveHost;
    const Xyzz<F> P = lazy_point<C>(a, za, ka), Q = lazy_point<C>(b, zb, kb);
    const RowConsts<W, F> K = row_consts<W, F>();
    const XyzzRow<W, F> rp = row_from_xyzz<W, F>(P), rq = row_from_xyzz<W, F>(Q);
    const XyzzRow<W, F> r = op == 0 ? row_add<W, F>(rp, rq, K) : row_dbl<W, F>(rp, K);
    const bool ok = row_consistent<F>(r.X) && row_consistent<F>(r.Y) && row_consistent<F>(r.ZZ) &&
                    row_consistent<F>(r.ZZZ);
    const Xyzz<F> rr = row_to_xyzz<W, F>(r);
    Xyzz<F> qq;
    run_quad<F>([&] { return op == 0 ? xyzz_add_quad<F, MulCallLazy, HostQuad>(P, Q)
                                     : xyzz_dbl_quad<F, MulCallLazy, HostQuad>(P); }, qq);
    raw_xyzz<F>(rr, raw_row);
    raw_xyzz<F>(qq, raw_quad);
    write_wire<C>(xyzz_canon_acc<F>(rr), out);
    return ok ? 0 : -1;
}

// The window combine by rows: nwin window sums (wire points raised to lazy
// forms by z and k), Horner from the top with c doublings per window.
template <class C>
int row_combine(const uint8_t* wins, const uint8_t* zs, const int* ks, int nwin, int c, uint8_t* out) {
    using F = typename C::Fp;
    using W = WaveHost;
    const RowConsts<W, F> K = row_consts<W, F>();
    auto win = [&](int w) {
        return row_from_xyzz<W, F>(lazy_point<C>(wins + 2 * C::FB * w, zs + C::FB * w, ks + 4 * w));
    };
    XyzzRow<W, F> r = win(nwin - 1);
    for (int w = nwin - 2; w >= 0; --w) {
        if (!row_pt_is_inf<W, F>(r))
            for (int i = 0; i < c; ++i) r = row_dbl<W, F>(r, K);
        r = row_add<W, F>(r, win(w), K);
    }
    const bool ok = row_consistent<F>(r.X) && row_consistent<F>(r.Y) && row_consistent<F>(r.ZZ) &&
                    row_consistent<F>(r.ZZZ);
    write_wire<C>(xyzz_canon_acc<F>(row_to_xyzz<W, F>(r)), out);
    return ok ? 0 : -1;
}

// ---------------------------------------------------------------- serving
// A host executor for the serving scheduler (serve_core.h) with injected
// failures: it "commits" a request as the first 48 bytes of its evaluations
// (so the tests can tell results apart), flags a batch whose evaluations hold
// an element >= 2^255 (above r), throws from launch() for chosen batches
// (a staging allocation or launch failure) and reports a device failure from
// poll() for others.  Batches take `latency` to finish.
struct FakeExec {
    struct L {
        int curve = 0;
        int batch = -1;
        bool err = false, dev_fail = false;
        std::vector<std::vector<uint8_t>> results;
        std::chrono::steady_clock::time_point ready;
    };
    std::vector<L> lane;
    std::vector<int> launch_fail, device_fail;  // batch indices
    std::chrono::microseconds latency{200};
    std::mutex mu;
    int next_batch = 0;
    std::map<const uint8_t*, int> batch_of;  // request evaluations -> last batch that ran them

    explicit FakeExec(size_t lanes) : lane(lanes) {}
    size_t lanes() const { return lane.size(); }
    uint64_t footprint(int, const std::vector<uint32_t>& sizes) {
        uint64_t b = 0;
        for (uint32_t n : sizes) b += 32ull * n;
        return b;
    }
    void launch(size_t li, int curve, const std::vector<std::pair<const uint8_t*, uint32_t>>& src,
                const std::vector<uint32_t>&) {
        std::lock_guard<std::mutex> g(mu);
        const int b = next_batch++;
        for (auto& p : src) batch_of[p.first] = b;
        if (std::find(launch_fail.begin(), launch_fail.end(), b) != launch_fail.end())
            throw std::runtime_error("injected launch failure");
        L& l = lane[li];
        l.curve = curve;
        l.batch = b;
        l.dev_fail = std::find(device_fail.begin(), device_fail.end(), b) != device_fail.end();
        l.err = false;
        l.results.clear();
        for (auto& p : src) {
            for (uint32_t e = 0; e < p.second; ++e) l.err |= p.first[32 * e + 31] >= 0x80;
            l.results.emplace_back(p.first, p.first + 48);
        }
        l.ready = std::chrono::steady_clock::now() + latency;
    }
    int poll(size_t li) {
        std::lock_guard<std::mutex> g(mu);
        const L& l = lane[li];
        if (std::chrono::steady_clock::now() < l.ready) return 0;
        return l.dev_fail ? -1 : 1;
    }
    bool error_flag(size_t li, size_t) {
        std::lock_guard<std::mutex> g(mu);
        return lane[li].err;
    }
    const uint8_t* result(size_t li, size_t k) {
        std::lock_guard<std::mutex> g(mu);
        return lane[li].results[k].data();
    }
};

}  // namespace

// One NTT pass of L = 2^logL points with the device kernel's lazy butterflies
// (ntt_math.h) in its order: bit-reversed load, a lone twiddle-free stage 0
// for odd logL, then radix-4 stage pairs; the end as in a last pass
// (canonical; n^-1 folded in for the inverse).  in: L 32-byte LE integers
// below 2^256 (a pass input may be up to 2r); tw: w_L^e for e < L/2 and ninv,
// canonical 32-byte LE.  *max_top gets the largest top limb held between stages.
template <class F>
void ntt_lazy_pass(int logL, int inverse, const uint8_t* in, const uint8_t* tw, const uint8_t* ninv, uint8_t* out,
                   uint32_t* max_top) {
    const uint32_t L = 1u << logL;
    auto load = [](const uint8_t* b) {
        uint32_t w[8];
        for (int k = 0; k < 8; ++k) w[k] = (uint32_t)b[4 * k] | ((uint32_t)b[4 * k + 1] << 8) |
                                           ((uint32_t)b[4 * k + 2] << 16) | ((uint32_t)b[4 * k + 3] << 24);
        return fe_from_words<F>(w, 8);
    };
    std::vector<Fe<F>> x(L), w(std::max<uint32_t>(1, L / 2));
    for (uint32_t j = 0; j < L; ++j) {
        uint32_t r = 0;
        for (int b = 0; b < logL; ++b) r |= ((j >> b) & 1u) << (logL - 1 - b);
        x[r] = load(in + 32 * j);
    }
    for (size_t e = 0; e < w.size(); ++e) w[e] = fe_to_mont<F>(load(tw + 32 * e));
    uint32_t top = 0;
    auto note = [&] {
        for (auto& v : x) top = std::max(top, v.v[F::N - 1]);
    };
    uint32_t s = 0;
    if (logL & 1) {
        for (uint32_t p = 0; p < L / 2; ++p) ntt_bfly0<F>(x[2 * p], x[2 * p + 1]);
        note();
        s = 1;
    }
    for (; s < (uint32_t)logL; s += 2) {
        const uint32_t h = 1u << s, sh1 = logL - 1 - s, sh2 = logL - 2 - s;
        for (uint32_t q = 0; q < L / 4; ++q) {
            const uint32_t i = q & (h - 1), e = ((q >> s) << (s + 2)) + i;
            if (s == 0) ntt_radix4<F, true>(x[e], x[e + h], x[e + 2 * h], x[e + 3 * h], w[i << sh2], w[i << sh2],
                                            w[(i + h) << sh2]);
            else ntt_radix4<F, false>(x[e], x[e + h], x[e + 2 * h], x[e + 3 * h], w[i << sh1], w[i << sh2],
                                      w[(i + h) << sh2]);
        }
        note();
    }
    const Fe<F> nv = fe_to_mont<F>(load(ninv));
    for (uint32_t k = 0; k < L; ++k) {
        const Fe<F> y = inverse ? fe_mul<F>(x[k], nv) : fe_reduce_wide<F>(x[k]);
        uint32_t wd[8];
        fe_to_words<F>(y, wd, 8);
        for (int b = 0; b < 32; ++b) out[32 * k + b] = (uint8_t)(wd[b / 4] >> (8 * (b % 4)));
    }
    *max_top = top;
}

template <class F>
void reduce_wide_bytes(const uint8_t* in36, uint8_t* out32) {
    uint32_t w[9];
    for (int k = 0; k < 9; ++k) w[k] = (uint32_t)in36[4 * k] | ((uint32_t)in36[4 * k + 1] << 8) |
                                       ((uint32_t)in36[4 * k + 2] << 16) | ((uint32_t)in36[4 * k + 3] << 24);
    const Fe<F> y = fe_reduce_wide<F>(fe_from_words<F>(w, 9));
    uint32_t wd[8];
    fe_to_words<F>(y, wd, 8);
    for (int b = 0; b < 32; ++b) out32[b] = (uint8_t)(wd[b / 4] >> (8 * (b % 4)));
}

// ---------------------------------------------------------------- Edwards
// Banderwagon point operations of curve_ed.h.  Points arrive as affine wire
// (BE x || y) with a projective scale z (LE, non-zero): X = x z, Y = y z,
// T = x y z, Z = z.  Results leave as affine wire; -1 if the extended
// coordinates are inconsistent (T Z != X Y), 0 if the four quad lanes
// disagree, 1 otherwise.
namespace {

using BwF = BwFp;

Fe<BwF> bw_be(const uint8_t* b) { return fe_to_mont<BwF>(fe_from_be_bytes<BwF>(b, 32)); }

EdExt<BwF> bw_in(const uint8_t* w, const uint8_t* z_le) {
    const Fe<BwF> x = bw_be(w), y = bw_be(w + 32), z = fe_to_mont<BwF>(from_le<BwF>(z_le, 32));
    EdExt<BwF> p;
    p.X = fe_mul<BwF>(x, z);
    p.Y = fe_mul<BwF>(y, z);
    p.T = fe_mul<BwF>(fe_mul<BwF>(x, y), z);
    p.Z = z;
    return p;
}

int bw_out(const EdExt<BwF>& p, uint8_t* w) {
    if (!fe_eq<BwF>(fe_mul<BwF>(p.T, p.Z), fe_mul<BwF>(p.X, p.Y))) return -1;
    const Fe<BwF> zi = fe_inv<BwF>(p.Z);
    fe_to_be_bytes<BwF>(fe_from_mont<BwF>(fe_mul<BwF>(p.X, zi)), w, 32);
    fe_to_be_bytes<BwF>(fe_from_mont<BwF>(fe_mul<BwF>(p.Y, zi)), w + 32, 32);
    return 1;
}

// op: 0 add, 1 dbl, 2 mixed add (b affine, record t = d x y), 3 add_quad,
// 4 dbl_quad, 5 negate a then add b
int bw_op(int op, const uint8_t* a, const uint8_t* b, const uint8_t* za, const uint8_t* zb, uint8_t* out) {
    const EdExt<BwF> P = bw_in(a, za), Q = bw_in(b, zb);
    EdExt<BwF> r;
    bool same = true;
    switch (op) {
        case 0: r = ed_add<BwF>(P, Q); break;
        case 1: r = ed_dbl<BwF>(P); break;
        case 2: {
            EdAff<BwF> q;
            q.x = bw_be(b);
            q.y = bw_be(b + 32);
            q.t = fe_mul<BwF>(fe_mul<BwF>(q.x, q.y), fe_const<BwF>(EdConst::D));
            r = P;
            ed_madd<BwF>(r, q);
            break;
        }
        case 3: same = run_quad<BwF>([&] { return ed_add_quad<BwF, MulCall, HostQuad>(P, Q); }, r); break;
        case 4: same = run_quad<BwF>([&] { return ed_dbl_quad<BwF, MulCall, HostQuad>(P); }, r); break;
        default: r = ed_add<BwF>(ed_neg<BwF>(P), Q); break;
    }
    if (!same) return 0;
    return bw_out(r, out);
}

// The accumulation loop's lazily reduced mixed additions: a (scaled by z,
// every coordinate raised by q, the largest representative the loop accepts)
// plus the n affine records of pts, with no canonical step in between.
int bw_madd_chain(const uint8_t* a, const uint8_t* za, const uint8_t* pts, int n, uint8_t* out) {
    EdExt<BwF> P = bw_in(a, za);
    P.X = add_kp<BwF>(P.X, 1);
    P.Y = add_kp<BwF>(P.Y, 1);
    P.T = add_kp<BwF>(P.T, 1);
    P.Z = add_kp<BwF>(P.Z, 1);
    for (int i = 0; i < n; ++i) {
        EdAff<BwF> q;
        q.x = bw_be(pts + 64 * i);
        q.y = bw_be(pts + 64 * i + 32);
        q.t = fe_mul<BwF>(fe_mul<BwF>(q.x, q.y), fe_const<BwF>(EdConst::D));
        ed_madd_lazy<BwF>(P, q);
        const Fe<BwF>* c[4] = {&P.X, &P.Y, &P.T, &P.Z};
        for (const Fe<BwF>* v : c)  // the bound the loop relies on: below 2q
            if (!fe_gt<BwF>(add_kp<BwF>(fe_zero<BwF>(), 2), *v)) return -2;
    }
    return bw_out(ed_canon<BwF>(P), out);
}

bool bw_below_2q(const EdExt<BwF>& p) {
    const Fe<BwF> two_q = add_kp<BwF>(fe_zero<BwF>(), 2);
    return fe_gt<BwF>(two_q, p.X) && fe_gt<BwF>(two_q, p.Y) && fe_gt<BwF>(two_q, p.T) && fe_gt<BwF>(two_q, p.Z);
}

EdExt<BwF> bw_raise(EdExt<BwF> p) {  // the largest representative below 2q
    p.X = add_kp<BwF>(p.X, 1);
    p.Y = add_kp<BwF>(p.Y, 1);
    p.T = add_kp<BwF>(p.T, 1);
    p.Z = add_kp<BwF>(p.Z, 1);
    return p;
}

// The lazily reduced quad forms over four host lanes, in a chain with no
// canonical step: `rounds` times P = 2P; P = P + Q, from the largest
// representatives of a and b; -2 if a coordinate ever reaches 2q, 0 if the
// lanes disagree.
int bw_quad_lazy_chain(const uint8_t* a, const uint8_t* b, const uint8_t* za, const uint8_t* zb, int rounds,
                       int inline_mul, uint8_t* out) {
    EdExt<BwF> P = bw_raise(bw_in(a, za));
    const EdExt<BwF> Q = bw_raise(bw_in(b, zb));
    for (int i = 0; i < rounds; ++i) {
        EdExt<BwF> r;
        bool same = inline_mul
            ? run_quad<BwF>([&] { return ed_dbl_quad_lazy<BwF, MulInline, MulLazy, HostQuad>(P); }, r)
            : run_quad<BwF>([&] { return ed_dbl_quad_lazy<BwF, MulCall, MulCallLazy, HostQuad>(P); }, r);
        if (!same) return 0;
        if (!bw_below_2q(r)) return -2;
        P = r;
        same = inline_mul
            ? run_quad<BwF>([&] { return ed_add_quad_lazy<BwF, MulInline, MulLazy, HostQuad>(P, Q); }, r)
            : run_quad<BwF>([&] { return ed_add_quad_lazy<BwF, MulCall, MulCallLazy, HostQuad>(P, Q); }, r);
        if (!same) return 0;
        if (!bw_below_2q(r)) return -2;
        P = r;
    }
    return bw_out(ed_canon<BwF>(P), out);
}

// Euler's criterion as the Banderwagon reader applies it.
int bw_is_square(const uint8_t* a_be) { return fe_is_square_nonzero<BwF>(bw_be(a_be)) ? 1 : 0; }

}  // namespace

// The row-distributed Edwards operations (field_row.h) in the chain of
// bw_quad_lazy_chain: every step's coordinates against the lazy quad forms,
// limb for limb; 0 when they all agree, then the canonical result in out.
int bw_row_chain(const uint8_t* a, const uint8_t* b, const uint8_t* za, const uint8_t* zb, int rounds,
                 uint8_t* out) {
    using W = WaveHost;
    EdExt<BwF> P = bw_raise(bw_in(a, za));
    const EdExt<BwF> Q = bw_raise(bw_in(b, zb));
    const RowConsts<W, BwF> K = row_consts<W, BwF>();
    const FeRow<W, BwF> d = row_from_fe<W, BwF>(fe_const<BwF>(EdConst::D));
    EdRow<W, BwF> R = row_from_ed<W, BwF>(P);
    const EdRow<W, BwF> RQ = row_from_ed<W, BwF>(Q);
    auto same = [](const EdExt<BwF>& x, const EdRow<W, BwF>& y) {
        const EdExt<BwF> z = row_to_ed<W, BwF>(y);
        return std::memcmp(&x, &z, sizeof x) == 0 && row_consistent<BwF>(y.X) && row_consistent<BwF>(y.Y) &&
               row_consistent<BwF>(y.T) && row_consistent<BwF>(y.Z);
    };
    for (int i = 0; i < rounds; ++i) {
        EdExt<BwF> r;
        run_quad<BwF>([&] { return ed_dbl_quad_lazy<BwF, MulCall, MulCallLazy, HostQuad>(P); }, r);
        P = r;
        R = row_ed_dbl<W, BwF>(R, K);
        if (!same(P, R)) return -1;
        run_quad<BwF>([&] { return ed_add_quad_lazy<BwF, MulCall, MulCallLazy, HostQuad>(P, Q); }, r);
        P = r;
        R = row_ed_add<W, BwF>(R, RQ, d, K);
        if (!same(P, R)) return -2;
    }
    return bw_out(ed_canon<BwF>(row_to_ed<W, BwF>(R)), out);
}

extern "C" {
int hc_bw_op(int op, const uint8_t* a, const uint8_t* b, const uint8_t* za, const uint8_t* zb, uint8_t* out) {
    return bw_op(op, a, b, za, zb, out);
}
int hc_bw_row_chain(const uint8_t* a, const uint8_t* b, const uint8_t* za, const uint8_t* zb, int rounds,
                    uint8_t* out) {
    return bw_row_chain(a, b, za, zb, rounds, out);
}
int hc_bw_is_square(const uint8_t* a_be) { return bw_is_square(a_be); }
int hc_bw_quad_lazy_chain(const uint8_t* a, const uint8_t* b, const uint8_t* za, const uint8_t* zb, int rounds,
                          int inline_mul, uint8_t* out) {
    return bw_quad_lazy_chain(a, b, za, zb, rounds, inline_mul, out);
}
int hc_bw_madd_chain(const uint8_t* a, const uint8_t* za, const uint8_t* pts, int n, uint8_t* out) {
    return bw_madd_chain(a, za, pts, n, out);
}
}  // extern "C"

extern "C" {

// One serving scenario on the host scheduler with the fake executor: nreq
// requests (curve[i], n[i] elements, bad[i] != 0 puts an element >= r into
// it) are submitted at once, then waited for.  launch_fail / device_fail:
// batch indices (in launch order) whose launch throws / whose device work
// fails, each list ended by -1.  Outputs: status[i] (0 or the error code),
// the first 8 bytes of each commitment (commit8[i]), the batch that ran each
// request last (batch[i]) and the 9 statistics in kzg_server_stats order.
int hc_serve_scenario(uint32_t lanes, uint32_t max_batch, uint64_t max_batch_elems, uint64_t budget, size_t nreq,
                      const int* curve, const uint32_t* n, const int* bad, const int* launch_fail,
                      const int* device_fail, int latency_us, int* status, uint64_t* commit8, int* batch,
                      uint64_t* stats_out) {
    FakeExec ex(lanes);
    for (const int* p = launch_fail; p && *p >= 0; ++p) ex.launch_fail.push_back(*p);
    for (const int* p = device_fail; p && *p >= 0; ++p) ex.device_fail.push_back(*p);
    ex.latency = std::chrono::microseconds(latency_us);
    std::vector<std::vector<uint8_t>> ev(nreq);
    uint32_t maxn = 0;
    for (size_t i = 0; i < nreq; ++i) {
        ev[i].assign(32 * (size_t)std::max<uint32_t>(n[i], 2), 0);
        for (int k = 0; k < 8; ++k) ev[i][k] = (uint8_t)((i + 1) >> (8 * k));  // the request's id
        if (bad[i]) ev[i][32 * (n[i] - 1) + 31] = 0xFF;
        maxn = std::max(maxn, n[i]);
    }
    kzg_server_config cfg{lanes, max_batch, max_batch_elems, budget};
    const size_t max_n[3] = {0, maxn, maxn};
    std::vector<int64_t> tk(nreq);
    kzg_server_stats st;
    {
        ServeCore<FakeExec> core(ex, cfg, max_n);
        for (size_t i = 0; i < nreq; ++i) tk[i] = core.submit((uint32_t)curve[i], ev[i].data(), n[i]);
        for (size_t i = 0; i < nreq; ++i) {
            uint8_t c[48] = {0};
            status[i] = tk[i] < 0 ? (int)tk[i] : core.wait(tk[i], c, nullptr, nullptr);
            uint64_t v = 0;
            std::memcpy(&v, c, 8);
            commit8[i] = v;
        }
        st = core.stats();
    }
    for (size_t i = 0; i < nreq; ++i) batch[i] = ex.batch_of.count(ev[i].data()) ? ex.batch_of[ev[i].data()] : -1;
    const uint64_t sv[9] = {st.submitted, st.completed, st.failed, st.batches, st.max_batch_requests,
                            st.peak_inflight_bytes, st.budget_waits, st.rerun_batches, st.launch_failures};
    std::memcpy(stats_out, sv, sizeof sv);
    return 0;
}

// The device's quad-cooperative point operations run on the host (four
// lock-stepped threads as the quad's lanes).  curve 0 secp256k1, 1 BN254,
// 2 BLS12-381; op 0 a + b, 1 2a; lazy 0 canonical products, 1 lazy; z: one
// LE coordinate-size value per point (nonzero), k: four coordinate offsets
// per point (multiples of p added to X, Y, ZZ, ZZZ).  Return 0 when the four
// lanes agreed.
int hc_quad_op(int curve, int op, int lazy, const uint8_t* a, const uint8_t* b, const uint8_t* za, const uint8_t* zb,
               const int* ka, const int* kb, uint8_t* out) {
    if (curve == 2) return quad_op<CurveBLS12381>(op, lazy, a, b, za, zb, ka, kb, out);
    if (curve == 1) return quad_op<CurveBN254>(op, lazy, a, b, za, zb, ka, kb, out);
    return quad_op<CurveSecp256k1>(op, lazy, a, b, za, zb, ka, kb, out);
}

int hc_quad_tree(int curve, int lazy, const uint8_t* pts, const uint8_t* zs, int n, uint8_t* out) {
    if (curve == 2) return quad_tree<CurveBLS12381>(lazy, pts, zs, n, out);
    if (curve == 1) return quad_tree<CurveBN254>(lazy, pts, zs, n, out);
    return quad_tree<CurveSecp256k1>(lazy, pts, zs, n, out);
}

// curve: 0 secp256k1, 1 BN254, 2 BLS12-381 (their Fp)
int hc_row_mul(int curve, const uint32_t* a, const uint32_t* b, uint32_t* out_row, uint32_t* out_ref) {
    if (curve == 2) return row_mul_check<BlsFp>(a, b, out_row, out_ref);
    if (curve == 1) return row_mul_check<BnFp>(a, b, out_row, out_ref);
    return row_mul_check<SecpFp>(a, b, out_row, out_ref);
}

int hc_row_op(int curve, int op, const uint8_t* a, const uint8_t* b, const uint8_t* za, const uint8_t* zb,
              const int* ka, const int* kb, uint8_t* out, uint32_t* raw_row, uint32_t* raw_quad) {
    if (curve == 2) return row_op<CurveBLS12381>(op, a, b, za, zb, ka, kb, out, raw_row, raw_quad);
    if (curve == 1) return row_op<CurveBN254>(op, a, b, za, zb, ka, kb, out, raw_row, raw_quad);
    return row_op<CurveSecp256k1>(op, a, b, za, zb, ka, kb, out, raw_row, raw_quad);
}

int hc_row_combine(int curve, const uint8_t* wins, const uint8_t* zs, const int* ks, int nwin, int c, uint8_t* out) {
    if (curve == 2) return row_combine<CurveBLS12381>(wins, zs, ks, nwin, c, out);
    if (curve == 1) return row_combine<CurveBN254>(wins, zs, ks, nwin, c, out);
    return row_combine<CurveSecp256k1>(wins, zs, ks, nwin, c, out);
}

int hc_quad_dbl_chain(int curve, const uint8_t* a, const uint8_t* za, const int* ka, int count, uint8_t* out) {
    if (curve == 2) return quad_dbl_chain<CurveBLS12381>(a, za, ka, count, out);
    if (curve == 1) return quad_dbl_chain<CurveBN254>(a, za, ka, count, out);
    return quad_dbl_chain<CurveSecp256k1>(a, za, ka, count, out);
}

// field: 0 BLS12-381 Fr, 1 BN254 Fr
void hc_ntt_lazy_pass(int field, int logL, int inverse, const uint8_t* in, const uint8_t* tw, const uint8_t* ninv,
                      uint8_t* out, uint32_t* max_top) {
    if (field == 0) ntt_lazy_pass<BlsFr>(logL, inverse, in, tw, ninv, out, max_top);
    else ntt_lazy_pass<BnFr>(logL, inverse, in, tw, ninv, out, max_top);
}

// x mod r for x < 2^261 given as 36 LE bytes (fe_reduce_wide: x below 64r)
void hc_reduce_wide(int field, const uint8_t* in36, uint8_t* out32) {
    if (field == 0) reduce_wide_bytes<BlsFr>(in36, out32);
    else reduce_wide_bytes<BnFr>(in36, out32);
}

// field: 0 BLS Fp, 1 BLS Fr, 2 BN Fp, 3 BN Fr, 4 secp256k1 Fp
void hc_field_op(int field, int op, const uint8_t* a, const uint8_t* b, uint8_t* out) {
    switch (field) {
        case 0: field_op<BlsFp>(op, a, b, out, 48); break;
        case 1: field_op<BlsFr>(op, a, b, out, 32); break;
        case 2: field_op<BnFp>(op, a, b, out, 32); break;
        case 4: field_op<SecpFp>(op, a, b, out, 32); break;
        default: field_op<BnFr>(op, a, b, out, 32); break;
    }
}

// BN254 GLV split of a 32-byte LE scalar: out = |k1| (16 B LE) || |k2| (16 B)
// || sign of k1 || sign of k2
void hc_glv_split_bn(const uint8_t* s32, uint8_t* out) {
    uint32_t s[8] = {0}, k1[4], k2[4];
    for (int k = 0; k < 32; ++k) s[k / 4] |= (uint32_t)s32[k] << (8 * (k % 4));
    bool n1, n2;
    glvbn::split(s, k1, k2, n1, n2);
    for (int k = 0; k < 16; ++k) {
        out[k] = (uint8_t)(k1[k / 4] >> (8 * (k % 4)));
        out[16 + k] = (uint8_t)(k2[k / 4] >> (8 * (k % 4)));
    }
    out[32] = n1 ? 1 : 0;
    out[33] = n2 ? 1 : 0;
}

void hc_glv_split_secp(const uint8_t* s32, uint8_t* out) {
    uint32_t s[8] = {0}, k1[4], k2[4];
    for (int k = 0; k < 32; ++k) s[k / 4] |= (uint32_t)s32[k] << (8 * (k % 4));
    bool n1, n2;
    glvsecp::split(s, k1, k2, n1, n2);
    for (int k = 0; k < 16; ++k) {
        out[k] = (uint8_t)(k1[k / 4] >> (8 * (k % 4)));
        out[16 + k] = (uint8_t)(k2[k / 4] >> (8 * (k % 4)));
    }
    out[32] = n1 ? 1 : 0;
    out[33] = n2 ? 1 : 0;
}

void hc_glv_split_bls(const uint8_t* s32, uint8_t* out) {
    uint32_t s[8] = {0}, k1[4], k2[4];
    for (int k = 0; k < 32; ++k) s[k / 4] |= (uint32_t)s32[k] << (8 * (k % 4));
    bool n1, n2;
    glvbls::split(s, k1, k2, n1, n2);
    for (int k = 0; k < 16; ++k) {
        out[k] = (uint8_t)(k1[k / 4] >> (8 * (k % 4)));
        out[16 + k] = (uint8_t)(k2[k / 4] >> (8 * (k % 4)));
    }
    out[32] = n1 ? 1 : 0;
    out[33] = n2 ? 1 : 0;
}

void hc_point_op(int curve, int op, const uint8_t* a, const uint8_t* b, uint8_t* out) {
    if (curve == 2) point_op<CurveBLS12381>(op, a, b, out);
    else if (curve == 1) point_op<CurveBN254>(op, a, b, out);
    else point_op<CurveSecp256k1>(op, a, b, out);
}

}  // extern "C"
