// resample.hip -- replication counts, rep -> index and in-place gather.
//
// Reference: include/vsmc/resample/transform.hpp (resample_trans_*),
// include/vsmc/resample/{multinomial,stratified,systematic,residual,
// residual_stratified,residual_systematic}.hpp, include/vsmc/rng/
// u01_sequence.hpp (the uniform sequences), include/vsmc/core/
// state_matrix.hpp:246-262 (copy).
//
// Counts: with s_i the sequential cumulative weight (exact, fold.hpp) and
// v_j the running maximum of the uniforms, the merge loop of
// transform.hpp:61-70 leaves j = G(s_i) = #{j : v_j <= s_i} after particle
// i, so rep_i = G(s_i) - G(s_{i-1}) and rep_{M-1} = N - G(s_{M-2}).
// G comes from closed forms for systematic and stratified uniforms and from
// a search over the materialised sequence otherwise.
#include "internal.hpp"
#include "particle.hpp"
#include "../../include/vsmch/device_rng.hpp"

#include <climits>
#include <mutex>
#include <type_traits>
#include <vector>

namespace vsmch {

using fold::FMap;
using fold::TileSeg;

// ------------------------------------------------------------ sources

struct WArray {
    const double *w;
    __device__ double operator()(long long i) const { return w[i]; }
    __device__ long long integ(long long) const { return 0; }
    // the source with its event parameters in registers (see WLog)
    __device__ WArray loaded() const { return *this; }
};

// WLog with its parameters read once: no per-element parameter loads or
// branches, so a tile's loads can all be in flight at once.
struct WLogV {
    const double *l;
    double inv_sum;
    double lmax;
    double eqw;
    int equal;
    const double *tab;   // kWexpTab or a copy in LDS
    __device__ double operator()(long long i) const { return from_raw(raw(i)); }
    // the log-weight alone (so that a tile's loads can be issued together)
    // and the weight from it
    __device__ double raw(long long i) const { return l[i]; }
    __device__ double from_raw(double li) const
    {
        const double v = inv_sum * wexp(li - lmax, tab);
        return equal ? eqw : v;
    }
    __device__ long long integ(long long) const { return 0; }
};

// W_i = inv_sum * e^(l_i - lmax) (Weight::normalize, weight.hpp:208-222;
// e^x by wexp)
struct WLog {
    const double *l;
    const ResampleDev *p;
    __device__ double operator()(long long i) const
    {
        if (p->equal)
            return 1.0 / static_cast<double>(p->M);
        return p->inv_sum * wexp(l[i] - p->lmax);
    }
    __device__ long long integ(long long) const { return 0; }
    __device__ WLogV loaded() const
    {
        return WLogV{l, p->inv_sum, p->lmax, 1.0 / static_cast<double>(p->M),
            p->equal, kWexpTab};
    }
};

// resid_i = modf(N * W_i) (transform.hpp:159-161)
template <class B>
struct WResidRaw {
    B b;
    double coeff;
    __device__ double operator()(long long i) const
    {
        double ip;
        return modf(coeff * b(i), &ip);
    }
    __device__ long long integ(long long i) const
    {
        double ip;
        modf(coeff * b(i), &ip);
        return static_cast<long long>(ip);
    }
    __device__ auto loaded() const
    {
        using BV = decltype(b.loaded());
        return WResidRaw<BV>{b.loaded(), coeff};
    }
};

// WResid with its parameters read once (see WLogV).
template <class B>
struct WResidV {
    B b;
    double coeff;
    double inv_resid;
    int active;
    __device__ double operator()(long long i) const
    {
        double ip;
        const double r = modf(coeff * b(i), &ip);
        return active ? inv_resid * r : 0.0;
    }
    __device__ long long integ(long long i) const
    {
        double ip;
        modf(coeff * b(i), &ip);
        return static_cast<long long>(ip);
    }
};

// normalised residual inv_resid * resid_i (transform.hpp:163-164); zero
// when no residual draws remain.
template <class B>
struct WResid {
    B b;
    double coeff;
    const ResampleDev *p;
    __device__ double operator()(long long i) const
    {
        if (p->ndraw <= 0)
            return 0.0;
        double ip;
        const double r = modf(coeff * b(i), &ip);
        return p->inv_resid * r;
    }
    __device__ long long integ(long long i) const
    {
        double ip;
        modf(coeff * b(i), &ip);
        return static_cast<long long>(ip);
    }
    __device__ auto loaded() const
    {
        using BV = decltype(b.loaded());
        return WResidV<BV>{b.loaded(), coeff, p->inv_resid, p->ndraw > 0 ? 1 : 0};
    }
};

// Sequence terms of U01SequenceSorted (u01_sequence.hpp:149-151), negated:
// a_k = -(log(U_k) / (n - k)) >= 0, so the fold of a is -lmax.
struct WSortedTermsV {
    unsigned int key;
    unsigned long long word0;
    long long ndraw;
    __device__ double operator()(long long k) const
    {
        const double u = u01_u32d(philox_word(key, word0 + k));
        return -(log(u) / static_cast<double>(ndraw - k));
    }
    __device__ long long integ(long long) const { return 0; }
};
struct WSortedTerms {
    const ResampleDev *p;
    __device__ double operator()(long long k) const
    {
        const double u = u01_u32d(philox_word(p->key, p->word0 + k));
        return -(log(u) / static_cast<double>(p->ndraw - k));
    }
    __device__ long long integ(long long) const { return 0; }
    __device__ WSortedTermsV loaded() const
    {
        return WSortedTermsV{p->key, p->word0, p->ndraw};
    }
};

// U01SequenceSystematic: u_j = u0 + j * delta (u01_sequence.hpp:246), no FMA.
struct USys {
    const ResampleDev *p;
    __device__ double u(long long j) const
    {
        return rn_add(p->u0, rn_mul(static_cast<double>(j), p->delta));
    }
    __device__ double ui(int j) const
    {
        return rn_add(p->u0, rn_mul(static_cast<double>(j), p->delta));
    }
    __device__ long long ndraw() const { return p->ndraw; }
    // Number of draws u_j <= x. Draw counts are below 2^31 (particle
    // indices are int32), so the search runs on 32-bit integers.
    __device__ long long count(double x) const
    {
        const int n = static_cast<int>(p->ndraw);
        if (n <= 0)
            return 0;
        if (!(x < INFINITY))
            return n;
        const double est = (x - p->u0) * static_cast<double>(n);
        int g;
        if (!(est > 0.0))
            g = 0;
        else if (est >= static_cast<double>(n))
            g = n;
        else
            g = static_cast<int>(est) + 1;
        while (g > 0 && ui(g - 1) > x)
            --g;
        while (g < n && ui(g) <= x)
            ++g;
        return g;
    }
};

// U01SequenceStratified: u_j = U_j * delta + j * delta (u01_sequence.hpp:200),
// no FMA; v_j = max(u_{j-1}, u_j) is the running maximum because
// u_{j-2} < fl(j * delta) <= u_j.
struct UStrat {
    const ResampleDev *p;
    __device__ double u(long long j) const
    {
        const double U = u01_u32d(philox_word(p->key, p->word0 + j));
        return rn_add(rn_mul(U, p->delta),
            rn_mul(static_cast<double>(j), p->delta));
    }
    __device__ double v(long long j) const
    {
        const double a = u(j);
        if (j == 0)
            return a;
        const double b = u(j - 1);
        return b > a ? b : a;
    }
    __device__ double ui(int j) const
    {
        const double U = u01_u32d(philox_word(p->key,
            p->word0 + static_cast<unsigned long long>(j)));
        return rn_add(rn_mul(U, p->delta),
            rn_mul(static_cast<double>(j), p->delta));
    }
    __device__ double vi(int j) const
    {
        const double a = ui(j);
        if (j == 0)
            return a;
        const double b = ui(j - 1);
        return b > a ? b : a;
    }
    __device__ long long ndraw() const { return p->ndraw; }
    // Number of draws v_j <= x (32-bit search, as for USys). The search
    // starts at g = floor(x n) and almost always ends within one draw of
    // it, so the uniforms of draws g - 2 .. g + 1 are taken from the two
    // Philox blocks that hold them (one block evaluation per two words
    // instead of one per word); a search that leaves that window evaluates
    // the draws one at a time.
    __device__ long long count(double x) const
    {
        const int n = static_cast<int>(p->ndraw);
        if (n <= 0)
            return 0;
        if (!(x < INFINITY))
            return n;
        const double est = x * static_cast<double>(n);
        int g;
        if (!(est > 0.0))
            g = 0;
        else if (est >= static_cast<double>(n))
            g = n;
        else
            g = static_cast<int>(est);
        const int w0 = g >= 2 ? g - 2 : 0;
        const unsigned long long k0 = p->word0 + static_cast<unsigned long long>(w0);
        const u32x4 B0 = philox_block(p->key, k0 / 4 + 1);
        const u32x4 B1 = philox_block(p->key, k0 / 4 + 2);
        const unsigned off = static_cast<unsigned>(k0 & 3);
        // the uniforms of draws w0 .. w0 + 3: words off .. off + 3 of the
        // eight in B0, B1 (value selects: no indexed array in scratch)
        auto uwin = [&](unsigned q) -> double {
            const unsigned i = off + q;
            const unsigned r = i & 3;
            const uint32_t a = r == 0 ? B0.v[0] : r == 1 ? B0.v[1] : r == 2 ? B0.v[2] : B0.v[3];
            const uint32_t b = r == 0 ? B1.v[0] : r == 1 ? B1.v[1] : r == 2 ? B1.v[2] : B1.v[3];
            return rn_add(rn_mul(u01_u32d(i < 4 ? a : b), p->delta),
                rn_mul(static_cast<double>(w0 + static_cast<int>(q)), p->delta));
        };
        const double u0 = uwin(0), u1 = uwin(1), u2 = uwin(2), u3 = uwin(3);
        const double m1 = u0 > u1 ? u0 : u1;
        const double m2 = u1 > u2 ? u1 : u2;
        const double m3 = u2 > u3 ? u2 : u3;
        // v_j = max(u_{j-1}, u_j) (v_0 = u_0) for j in the window, any
        // other draw one at a time
        auto vget = [&](int j) -> double {
            const int q = j - w0;
            if (q == 0 && w0 == 0)
                return u0;
            if (q == 1)
                return m1;
            if (q == 2)
                return m2;
            if (q == 3)
                return m3;
            return vi(j);
        };
        while (g > 0 && vget(g - 1) > x)
            --g;
        while (g < n && vget(g) <= x)
            ++g;
        return g;
    }
};

// Materialised non-decreasing sequence v (running maximum of the uniforms).
struct UArr {
    const double *v;
    const ResampleDev *p;
    __device__ long long ndraw() const { return p->ndraw; }
    __device__ long long count(double x) const
    {
        long long lo = 0, hi = p->ndraw;
        while (lo < hi) {
            const long long mid = (lo + hi) >> 1;
            if (v[mid] <= x)
                lo = mid + 1;
            else
                hi = mid;
        }
        return lo;
    }
};

__device__ inline long long eff_n(const long long *ndev, long long nhost)
{
    return ndev ? *ndev : nhost;
}

__device__ inline long long ceil_div(long long a, long long b)
{
    return (a + b - 1) / b;
}

// ---------------------------------------------------------- fold passes

template <class WSrc>
__global__ void __launch_bounds__(kBlock) k_tile_sum(WSrc src, long long nh,
    const long long *ndev, const ResampleDev *prm, double *tsum, int *err)
{
    if (!prm->active)
        return;
    const long long n = eff_n(ndev, nh);
    const long long base = static_cast<long long>(blockIdx.x) * kTile;
    if (base >= n)
        return;
    __shared__ double lds[kWaves];
    double acc = 0.0;
    int bad = 0;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const long long g = base + j * kBlock + threadIdx.x;
        if (g < n) {
            const double x = src(g);
            if (!(x >= 0.0))
                bad = 1;
            acc += x;
        }
    }
    const double tot = block_reduce(acc, OpAddD(), lds);
    if (threadIdx.x == 0)
        tsum[blockIdx.x] = tot;
    if (bad)
        atomicOr(err, 1);
}

__global__ void __launch_bounds__(kBlock) k_scan_tiles(const double *tsum,
    long long nh, const long long *ndev, const ResampleDev *prm, double *tpre)
{
    if (!prm->active)
        return;
    const long long n = eff_n(ndev, nh);
    const long long ntiles = ceil_div(n, kTile);
    __shared__ double lds[kWaves];
    double carry = 0.0;
    for (long long c = 0; c < ntiles; c += kBlock) {
        const long long t = c + threadIdx.x;
        const double v = t < ntiles ? tsum[t] : 0.0;
        double tot;
        const double ex = block_exclusive_scan(v, OpAddD(), 0.0, lds, &tot);
        if (t < ntiles)
            tpre[t] = carry + ex;
        carry += tot;
    }
}

// Summary of group g (64 tiles), by one wave: flat when every tile is a
// single run in the same binade; the map is the composition of the tiles'.
__device__ __forceinline__ void group_summary(const TileSeg *segs,
    long long ntiles, long long g, GroupSeg *groups)
{
    const int lane = static_cast<int>(threadIdx.x % kWave);
    const long long t = g * kGroup + lane;
    const bool valid = t < ntiles;
    int nc = 0, e = fold::kEmpty;
    FMap m = fold::ident();
    if (valid) {
        nc = segs[t].ncross;
        e = segs[t].e[0];
        m = segs[t].map[0];
    }
    const int e0 = __shfl(e, 0, kWave);
    const bool ok = !valid || (nc == 0 && e == e0 && e != fold::kEmpty);
    const unsigned long long bal = __ballot(ok);
    const FMap P = wave_inclusive_scan(m, fold::OpCompose());
    const FMap tot = shfl_any(P, kWave - 1);
    if (lane == 0) {
        GroupSeg gs;
        gs.flat = (bal == ~0ull) ? 1 : 0;
        gs.e = e0;
        gs.map = tot;
        groups[g] = gs;
    }
}

template <class WSrc>
__global__ void __launch_bounds__(kBlock) k_segments(WSrc src, long long nh,
    const long long *ndev, const ResampleDev *prm, const double *tpre,
    TileSeg *segs)
{
    if (!prm->active)
        return;
    const long long n = eff_n(ndev, nh);
    const long long base = static_cast<long long>(blockIdx.x) * kTile;
    if (base >= n)
        return;
    __shared__ fold::SegShared sh;
    load_tile_lds(src, base, n, sh.tile, 0.0);
    const long long rem = n - base;
    const int nvalid = rem < kTile ? static_cast<int>(rem) : kTile;
    fold::tile_segments(nvalid, blockIdx.x == 0 ? 0.0 : tpre[blockIdx.x],
        &segs[blockIdx.x], sh);
}

// One wave per group of 64 tiles (group_summary).
__global__ void __launch_bounds__(kWave) k_groups(long long nh,
    const long long *ndev, const ResampleDev *prm, const TileSeg *segs,
    GroupSeg *groups)
{
    if (!prm->active)
        return;
    const long long ntiles = ceil_div(eff_n(ndev, nh), kTile);
    if (static_cast<long long>(blockIdx.x) * kGroup >= ntiles)
        return;
    group_summary(segs, ntiles, blockIdx.x, groups);
}

constexpr int kRGroups = kGroup;   // group summaries held per window
constexpr int kRSlots = 16;        // non-flat groups whose tile summaries are held
constexpr int kRSegs = 16;         // crossing tiles whose full summaries are held

struct TileSeg1 {
    int nc;
    int e;
    FMap m;
};

struct ResolveShared {
    double tile[kTileLds];
    fold::ExactShared ex;
    long long lscan[kWaves];
    // the window's group summaries; gslot: slot of a non-flat group's tile
    // summaries (kRSlots: staged when the walk reaches it)
    int gflat[kRGroups];
    int ge[kRGroups];
    FMap gmap[kRGroups];
    int gslot[kRGroups];
    int slotg[kRSlots];
    // single-run tile summaries per slot (slot kRSlots stages one group);
    // tseg: a crossing tile's full summary in seg[] (kRSegs: staged)
    int tnc[kRSlots + 1][kGroup];
    int te[kRSlots + 1][kGroup];
    FMap tmap[kRSlots + 1][kGroup];
    int tseg[kRSlots][kGroup];
    // segments (consecutive advanceable items in one binade): start flag and
    // the composed map from the segment's first item through each item
    int gst[kRGroups];
    FMap gP[kRGroups];
    int tst[kRSlots + 1][kGroup];
    FMap tP[kRSlots + 1][kGroup];
    long long segt[kRSegs];
    TileSeg seg[kRSegs + 1];
    // walk state
    double x;
    long long g;
    long long t;     // tile inside the non-flat group g; -1 at group level
    int status;      // 1: tile t needs the exact fold
};

// Orders LDS writes of this wave before its later LDS reads.
__device__ inline void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// Segment start flags and prefix maps of 64 items for one wave (lane k:
// item k): an item is advanceable when `ok`; a segment is a run of
// advanceable items in one binade, each other item a segment of its own.
__device__ inline void wave_segments(bool valid, bool ok, int e, FMap m,
    int *st, FMap *P)
{
    const int lane = threadIdx.x % kWave;
    ok = valid && ok;
    const int eprev = __shfl_up(e, 1, kWave);
    const int okprev = __shfl_up(ok ? 1 : 0, 1, kWave);
    const bool start = lane == 0 || !ok || !okprev || e != eprev;
    const fold::SegPair sp = wave_inclusive_scan(
        fold::SegPair{start ? 1LL : 0LL, ok ? m : fold::ident()}, fold::OpSeg());
    st[lane] = start ? 1 : 0;
    P[lane] = sp.m;
}

// Batched advance over up to 64 items (groups or tiles) whose summaries are
// single maps in binade binade(x): returns how many were applied, writes the
// start of each into starts[], and updates x. Called by wave 0 only.
template <class Get>
__device__ inline int wave_advance(double x, long long first, long long count,
    Get get, double *starts, int *modes, double *xnew)
{
    const int lane = threadIdx.x;
    const bool xpos = x > 0.0 && x < INFINITY;
    const int ex = xpos ? fold::binade(x) : fold::kEmpty;
    const long long it = first + lane;
    bool ok = false;
    FMap m = fold::ident();
    if (xpos && lane < count) {
        int e;
        FMap mm;
        if (get(it, e, mm) && e == ex) {
            ok = true;
            m = mm;
        }
    }
    const unsigned long long bal = __ballot(ok);
    const int L = (~bal == 0ull) ? kWave : (__ffsll(~bal) - 1);
    const FMap P = wave_inclusive_scan(m, fold::OpCompose());
    const long long a0 = xpos ? fold::mant(x) : 0;
    const long long a = a0 + ((a0 & 1) ? P.d1 : P.d0);
    const bool over = lane < L && a >= fold::kTwo53;
    const unsigned long long bo = __ballot(over);
    int L2 = L;
    if (bo != 0ull) {
        const int f = __ffsll(bo) - 1;
        L2 = f < L2 ? f : L2;
    }
    const double val = (lane < L2) ? fold::from_mant(a, ex) : 0.0;
    double prevv = shfl_up_any(val, 1);
    if (lane == 0)
        prevv = x;
    if (lane < L2) {
        starts[it] = prevv;
        if (modes)
            modes[it] = 1;
    }
    const double last = shfl_any(val, L2 > 0 ? L2 - 1 : 0);
    *xnew = L2 > 0 ? last : x;
    return L2;
}

// wave_advance from the first item of a segment, with the segment's prefix
// maps precomputed: item(it, ok, e, st, P). Same results as wave_advance.
// Called by wave 0 only.
template <class Item>
__device__ inline int wave_advance_seg(double x, long long first,
    long long count, Item item, double *starts, int *modes, double *xnew)
{
    const int lane = threadIdx.x;
    const bool xpos = x > 0.0 && x < INFINITY;
    const int ex = xpos ? fold::binade(x) : fold::kEmpty;
    const long long it = first + lane;
    bool ok = false;
    FMap P = fold::ident();
    if (xpos && lane < count) {
        bool iok, st;
        int e;
        item(it, iok, e, st, P);
        ok = iok && e == ex && (lane == 0 || !st);
    }
    const unsigned long long bal = __ballot(ok);
    const int L = (~bal == 0ull) ? kWave : (__ffsll(~bal) - 1);
    const long long a0 = xpos ? fold::mant(x) : 0;
    const long long a = a0 + ((a0 & 1) ? P.d1 : P.d0);
    const unsigned long long bo = __ballot(lane < L && a >= fold::kTwo53);
    int L2 = L;
    if (bo != 0ull) {
        const int f = __ffsll(bo) - 1;
        L2 = f < L2 ? f : L2;
    }
    const double val = (lane < L2) ? fold::from_mant(a, ex) : 0.0;
    double prevv = shfl_up_any(val, 1);
    if (lane == 0)
        prevv = x;
    if (lane < L2) {
        starts[it] = prevv;
        if (modes)
            modes[it] = 1;
    }
    const double last = shfl_any(val, L2 > 0 ? L2 - 1 : 0);
    *xnew = L2 > 0 ? last : x;
    return L2;
}

// Stages the summaries of groups [gw0, gw0 + nwg) into LDS: every group's
// (flat, binade, map), the single-run tile summaries of the first kRSlots
// non-flat groups and the full summaries of the first kRSegs crossing tiles
// among them. All threads call it; it ends with a barrier.
__device__ __forceinline__ void resolve_stage(ResolveShared &sh,
    const TileSeg *segs, const GroupSeg *groups, long long gw0, int nwg,
    long long ntiles)
{
    const int i = threadIdx.x;
    long long nonflat = 0;
    {
        GroupSeg gs{0, fold::kEmpty, fold::ident()};
        if (i < nwg) {
            gs = groups[gw0 + i];
            sh.gflat[i] = gs.flat;
            sh.ge[i] = gs.e;
            sh.gmap[i] = gs.map;
            nonflat = gs.flat ? 0 : 1;
        }
        if (i < kWave)
            wave_segments(i < nwg, gs.flat != 0, gs.e, gs.map, sh.gst, sh.gP);
    }
    long long nslot;
    const long long s = block_exclusive_scan(nonflat, OpAddL(), 0LL, sh.lscan,
        &nslot);
    if (i < nwg) {
        const bool held = nonflat && s < kRSlots;
        sh.gslot[i] = held ? static_cast<int>(s) : kRSlots;
        if (held)
            sh.slotg[s] = i;
    }
    __syncthreads();
    const int ns = nslot < kRSlots ? static_cast<int>(nslot) : kRSlots;
    constexpr int kPer = kRSlots * kGroup / kBlock;
    static_assert(kPer * kBlock == kRSlots * kGroup, "slot entries per thread");
    int nc[kPer];
    long long cnt = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int q = i * kPer + k;
        const int sl = q / kGroup, lane = q % kGroup;
        nc[k] = 0;
        if (sl < ns) {
            const long long tile = (gw0 + sh.slotg[sl]) * kGroup + lane;
            if (tile < ntiles) {
                const TileSeg &sg = segs[tile];
                nc[k] = sg.ncross;
                sh.tnc[sl][lane] = sg.ncross;
                sh.te[sl][lane] = sg.e[0];
                sh.tmap[sl][lane] = sg.map[0];
                cnt += nc[k] > 0;
            }
        }
    }
    long long nseg;
    long long si = block_exclusive_scan(cnt, OpAddL(), 0LL, sh.lscan, &nseg);
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int q = i * kPer + k;
        const int sl = q / kGroup, lane = q % kGroup;
        if (sl < ns) {
            int idx = kRSegs;
            if (nc[k] > 0) {
                if (si < kRSegs) {
                    idx = static_cast<int>(si);
                    sh.segt[si] = (gw0 + sh.slotg[sl]) * kGroup + lane;
                }
                ++si;
            }
            sh.tseg[sl][lane] = idx;
        }
    }
    __syncthreads();
    for (int sl = i / kWave; sl < ns; sl += kWaves) {
        const int lane = i % kWave;
        const long long tile = (gw0 + sh.slotg[sl]) * kGroup + lane;
        const bool valid = tile < ntiles;
        wave_segments(valid, valid && sh.tnc[sl][lane] == 0,
            valid ? sh.te[sl][lane] : 0, valid ? sh.tmap[sl][lane] : fold::ident(),
            sh.tst[sl], sh.tP[sl]);
    }
    const int nsg = nseg < kRSegs ? static_cast<int>(nseg) : kRSegs;
    constexpr int kWords = static_cast<int>(sizeof(TileSeg) / 4);
    for (int q = i; q < nsg * kWords; q += kBlock) {
        const int k = q / kWords, w = q % kWords;
        reinterpret_cast<int *>(&sh.seg[k])[w] =
            reinterpret_cast<const int *>(&segs[sh.segt[k]])[w];
    }
    __syncthreads();
}

// Wave 0 walks from the state in sh (x, g, t) over the groups up to gend,
// with every summary read from LDS: flat groups and single-run tiles 64 at
// a time, crossing tiles through their full summary. Stops at the end of
// the window (status 0) or at a tile whose summary does not apply
// (status 1); the state goes back to sh.
__device__ __forceinline__ void resolve_walk(ResolveShared &sh,
    const TileSeg *segs, long long gw0, long long gend, long long ntiles,
    double *xstart, double *xgstart, int *gmode)
{
    constexpr int kWords = static_cast<int>(sizeof(TileSeg) / 4);
    const int lane = threadIdx.x;
    double x = sh.x;
    long long g = sh.g, t = sh.t;
    int status = 0;
    while (true) {
        if (t < 0) {
            if (g >= gend)
                break;
            double xn;
            int L;
            if (sh.gst[g - gw0]) {
                auto item = [&](long long gg, bool &ok, int &e, bool &st, FMap &P) {
                    const int k = static_cast<int>(gg - gw0);
                    ok = sh.gflat[k] != 0;
                    e = sh.ge[k];
                    st = sh.gst[k] != 0;
                    P = sh.gP[k];
                };
                L = wave_advance_seg(x, g, gend - g, item, xgstart, gmode, &xn);
            } else {
                auto get = [&](long long gg, int &e, FMap &m) {
                    const int k = static_cast<int>(gg - gw0);
                    e = sh.ge[k];
                    m = sh.gmap[k];
                    return sh.gflat[k] != 0;
                };
                L = wave_advance(x, g, gend - g, get, xgstart, gmode, &xn);
            }
            x = xn;
            if (L > 0) {
                g += L;
                continue;
            }
            // group g walked tile by tile
            if (lane == 0) {
                gmode[g] = 0;
                xgstart[g] = x;
            }
            t = g * kGroup;
            if (sh.gslot[g - gw0] == kRSlots) {
                const long long tile = t + lane;
                const bool valid = tile < ntiles;
                TileSeg1 one{0, 0, fold::ident()};
                if (valid) {
                    const TileSeg &sg = segs[tile];
                    one = TileSeg1{sg.ncross, sg.e[0], sg.map[0]};
                    sh.tnc[kRSlots][lane] = one.nc;
                    sh.te[kRSlots][lane] = one.e;
                    sh.tmap[kRSlots][lane] = one.m;
                }
                wave_segments(valid, one.nc == 0, one.e, one.m, sh.tst[kRSlots],
                    sh.tP[kRSlots]);
                wave_lds_sync();
            }
        }
        const long long t0 = g * kGroup;
        const long long tend = (t0 + kGroup < ntiles) ? t0 + kGroup : ntiles;
        if (t >= tend) {
            ++g;
            t = -1;
            continue;
        }
        const int sl = sh.gslot[g - gw0];
        double xn;
        int L2;
        if (sh.tst[sl][t - t0]) {
            auto item = [&](long long tt, bool &ok, int &e, bool &st, FMap &P) {
                const int k = static_cast<int>(tt - t0);
                ok = sh.tnc[sl][k] == 0;
                e = sh.te[sl][k];
                st = sh.tst[sl][k] != 0;
                P = sh.tP[sl][k];
            };
            L2 = wave_advance_seg(x, t, tend - t, item, xstart,
                static_cast<int *>(nullptr), &xn);
        } else {
            auto get = [&](long long tt, int &e, FMap &m) {
                const int k = static_cast<int>(tt - t0);
                e = sh.te[sl][k];
                m = sh.tmap[sl][k];
                return sh.tnc[sl][k] == 0;
            };
            L2 = wave_advance(x, t, tend - t, get, xstart,
                static_cast<int *>(nullptr), &xn);
        }
        x = xn;
        if (L2 > 0) {
            t += L2;
            continue;
        }
        if (lane == 0)
            xstart[t] = x;
        const int k = static_cast<int>(t - t0);
        if (sh.tnc[sl][k] > 0) {
            const int idx = sl < kRSlots ? sh.tseg[sl][k] : kRSegs;
            if (idx == kRSegs) {
                const int *from = reinterpret_cast<const int *>(&segs[t]);
                int *to = reinterpret_cast<int *>(&sh.seg[kRSegs]);
                for (int q = lane; q < kWords; q += kWave)
                    to[q] = from[q];
                wave_lds_sync();
            }
            double y = x;
            if (fold::apply_segments(sh.seg[idx], y)) {
                x = y;
                ++t;
                continue;
            }
        }
        status = 1;
        break;
    }
    if (lane == 0) {
        sh.x = x;
        sh.g = g;
        sh.t = t;
        sh.status = status;
    }
}

// Exact start of every tile (xstart) and of every group (xgstart, gmode 1
// where the group was advanced whole and k_starts fills its tiles). One
// workgroup: per window of kRGroups groups, the summaries are staged in LDS,
// wave 0 walks them, and the whole block evaluates a tile exactly where a
// prediction fails.
template <class WSrc>
__device__ __forceinline__ void resolve_body(WSrc src, long long n,
    const TileSeg *segs, const GroupSeg *groups, double *xstart,
    double *xgstart, int *gmode, ResolveShared &sh)
{
    const long long ntiles = ceil_div(n, kTile);
    const long long ngroups = ceil_div(ntiles, kGroup);
    if (threadIdx.x == 0)
        sh.x = 0.0;
    for (long long gw0 = 0; gw0 < ngroups; gw0 += kRGroups) {
        const long long gend = (gw0 + kRGroups < ngroups) ? gw0 + kRGroups : ngroups;
        resolve_stage(sh, segs, groups, gw0, static_cast<int>(gend - gw0), ntiles);
        if (threadIdx.x == 0) {
            sh.g = gw0;
            sh.t = -1;
        }
        __syncthreads();
        while (true) {
            if (threadIdx.x < kWave)
                resolve_walk(sh, segs, gw0, gend, ntiles, xstart, xgstart, gmode);
            __syncthreads();
            if (sh.status == 0)
                break;
            const long long t = sh.t;
            const long long base = t * kTile;
            load_tile_lds(src, base, n, sh.tile, 0.0);
            fold::tile_exact(sh.tile, sh.x, sh.ex);
            const long long rem = n - base;
            const int last = (rem < kTile ? static_cast<int>(rem) : kTile) - 1;
            const double y = sh.tile[lds_pad(last)];
            __syncthreads();
            if (threadIdx.x == 0) {
                sh.x = y;
                sh.t = t + 1;
            }
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        xstart[ntiles] = sh.x;
        xgstart[ngroups] = sh.x;
    }
}

template <class WSrc>
__global__ void __launch_bounds__(kBlock) k_resolve(WSrc src, long long nh,
    const long long *ndev, const ResampleDev *prm, const TileSeg *segs,
    const GroupSeg *groups, double *xstart, double *xgstart, int *gmode)
{
    if (!prm->active)
        return;
    __shared__ ResolveShared sh;
    resolve_body(src, eff_n(ndev, nh), segs, groups, xstart, xgstart, gmode, sh);
}

// Exact tile starts inside the groups the resolver advanced whole (gmode):
// tile t0 + k starts at the group start mapped by the first k tile maps,
// all in the group's binade and below its top. One wave per group; the
// resolver wrote xstart for the tiles of every other group.
__device__ __forceinline__ void starts_group(long long g, long long ntiles,
    const TileSeg *segs, const double *xgstart, const int *gmode,
    double *xstart)
{
    const int lane = static_cast<int>(threadIdx.x % kWave);
    const long long t0 = g * kGroup;
    if (t0 >= ntiles || !gmode[g])
        return;
    const long long t = t0 + lane;
    FMap m = fold::ident();
    if (t < ntiles)
        m = segs[t].map[0];
    const FMap P = wave_inclusive_scan(m, fold::OpCompose());
    FMap E = shfl_up_any(P, 1);
    if (lane == 0)
        E = fold::ident();
    if (t < ntiles) {
        const double x = xgstart[g];
        const long long a0 = fold::mant(x);
        xstart[t] = fold::from_mant(a0 + ((a0 & 1) ? E.d1 : E.d0),
            fold::binade(x));
    }
}

__global__ void __launch_bounds__(kWave) k_starts(long long nh,
    const long long *ndev, const ResampleDev *prm, const TileSeg *segs,
    const double *xgstart, const int *gmode, double *xstart)
{
    if (!prm->active)
        return;
    const long long ntiles = ceil_div(eff_n(ndev, nh), kTile);
    starts_group(blockIdx.x, ntiles, segs, xgstart, gmode, xstart);
}

// Last index i in [0, n) with at(i) <= x for a non-decreasing at(), or -1:
// one whole wave, 64 probes a round (every lane returns the result; the loop
// bounds are wave-uniform).
template <typename At, typename X>
__device__ inline long long wave_last_le(At at, long long n, X x)
{
    const long long lane = threadIdx.x & (kWave - 1);
    long long lo = 0, hi = n;   // the first index with at(i) > x lies in [lo, hi]
    while (lo < hi) {
        const long long step = (hi - lo + kWave - 1) / kWave;
        const long long p = lo + lane * step;
        const bool le = p < hi && at(p) <= x;
        const long long c = __popcll(__ballot(le));
        if (c == 0)
            break;
        const long long nhi = lo + c * step;
        lo = lo + (c - 1) * step + 1;
        hi = nhi < hi ? nhi : hi;
    }
    return lo - 1;
}

// Draws of a tile's slice of materialised sorted uniforms held in LDS at a
// time, and the most chunks a tile walks that way.
constexpr int kVSlice = 4096;
constexpr int kVChunks = 4;

struct FoldBufs {
    const double *xstart;   // exact start of every tile, total at [ntiles]
};

template <bool kFull, class WSrc, class USrc>
__device__ __forceinline__ double count_tile(WSrc src, USrc us, long long M,
    long long tidx, double x0, int *rep, long long *tagg, int *hard,
    int hard_id = -1);

// Every tile with the full exact fold (the one-binade fast path first), a
// tile whose sums cross a binade in the same block.
template <class WSrc, class USrc>
__global__ void __launch_bounds__(kBlock) k_count_all(WSrc src, USrc us,
    long long nh, const long long *ndev, const ResampleDev *prm, FoldBufs fb,
    int *rep, long long *tagg)
{
    if (!prm->active)
        return;
    const long long M = eff_n(ndev, nh);
    if (static_cast<long long>(blockIdx.x) * kTile >= M)
        return;
    count_tile<true>(src, us, M, blockIdx.x, fb.xstart[blockIdx.x], rep, tagg,
        static_cast<int *>(nullptr));
}

// Replication counts (and residual integer parts) of tile `tidx` of M
// weights from the exact start x0 of the tile; per-tile rep -> index
// aggregates to tagg[3 tidx ..] when tagg is non-null. Returns the exact
// sum after the tile's last element (the next tile's start). Without kFull
// only a tile whose sums stay in x0's binade is done; another is appended
// to `hard` (as hard_id when that is set, else as tidx) and nothing is
// written.
template <bool kFull, class WSrc, class USrc>
__device__ __forceinline__ double count_tile(WSrc src, USrc us, long long M,
    long long tidx, double x0, int *rep, long long *tagg, int *hard,
    int hard_id)
{
    const long long base = tidx * kTile;
    __shared__ double tile[kTileLds];
    __shared__ long long gfirst;
    __shared__ double xend;
    __shared__ long long red[kWaves];
    load_tile_lds(src, base, M, tile, 0.0);
    if constexpr (kFull) {
        __shared__ fold::ExactShared ex;
        fold::tile_exact(tile, x0, ex);
    } else {
        __shared__ fold::SingleShared ssh;
        if (!fold::tile_single(tile, x0, ssh)) {
            if (threadIdx.x == 0)
                hard[1 + atomicAdd(hard, 1)] = hard_id >= 0 ? hard_id : static_cast<int>(tidx);
            return 0.0;
        }
    }
    const long long nd = us.ndraw();
    const int nvalid = M - base < kTile ? static_cast<int>(M - base) : kTile;
    // count before the tile's first element: that of the previous sum
    // (0 before particle 0: the merge starts at j = 0)
    // (materialised draws: found with the tile's slice below)
    if (!std::is_same<USrc, UArr>::value && threadIdx.x == 0)
        gfirst = (tidx == 0) ? 0 : us.count(x0);
    // each sum s_k is replaced by its count G(s_k); the last particle takes
    // every remaining draw
    if constexpr (std::is_same<USrc, UArr>::value) {
        // Materialised sorted draws: the tile's counts lie in [G(x0),
        // G(xend)], so only the slice v[G(x0), G(xend)) decides them; it is
        // staged in LDS when it fits and every sum searched there, else
        // searched in place (instead of every sum searching all draws).
        __shared__ long long grange[2];
        __shared__ double vs[kVSlice];
        const int wave = static_cast<int>(threadIdx.x) / kWave;
        const double xl = tile[lds_pad(nvalid - 1)];
        if (wave < 2) {
            const double x = wave == 0 ? x0 : xl;
            const long long g = (wave == 0 && tidx == 0) ? 0 :
                wave_last_le([&](long long q) { return us.v[q]; }, nd, x) + 1;
            if ((threadIdx.x & (kWave - 1)) == 0)
                grange[wave] = g;
        }
        if (threadIdx.x == 0)
            xend = xl;
        __syncthreads();
        const long long glo = grange[0];
        const long long ghi = grange[1] > glo ? grange[1] : glo;
        const long long m = ghi - glo;
        // the slice in chunks of kVSlice draws: a sum's count is in the
        // first chunk whose last draw exceeds it (else it is m). Dense
        // draws (at least one per two sums): every sum searches the chunk
        // (strided order, the same number of steps on every lane). Sparse
        // draws: each thread takes kItems consecutive sums (blocked order),
        // which are sorted, a binary search for the first of them in a
        // chunk, then short forward steps from the previous answer (a
        // binary search past 8 steps).
        const bool dense = 2 * m >= nvalid;
        // a slice of more than kVChunks chunks (weight concentrated in the
        // tile) is searched in place by every sum instead, so that no block
        // walks a long slice alone
        const bool in_lds = m <= static_cast<long long>(kVChunks) * kVSlice;
        const int kb = static_cast<int>(threadIdx.x) * kItems;
        long long cnt[kItems];
        unsigned open = 0;   // bit j: the count of sum j not found yet
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const int k = dense ? j * kBlock + static_cast<int>(threadIdx.x) : kb + j;
            cnt[j] = m;
            open |= k < nvalid ? 1u << j : 0u;
        }
        if (!in_lds) {
#pragma unroll
            for (int j = 0; j < kItems; ++j) {
                if ((open >> j) & 1u) {
                    const int k = dense ? j * kBlock + static_cast<int>(threadIdx.x) : kb + j;
                    const double sk = tile[lds_pad(k)];
                    long long lo = 0, hi = m;   // first slice index with v > sk
                    while (lo < hi) {
                        const long long mid = (lo + hi) >> 1;
                        if (us.v[glo + mid] <= sk)
                            lo = mid + 1;
                        else
                            hi = mid;
                    }
                    cnt[j] = lo;
                }
            }
            open = 0;
        }
        for (long long c0 = 0; c0 < (in_lds ? m : 0); c0 += kVSlice) {
            const int cn = m - c0 < kVSlice ? static_cast<int>(m - c0) : kVSlice;
            {
                // clamped, unconditional loads, all in flight together
                static_assert(kVSlice == kItems * kBlock, "one load per item");
                const double *const vc = us.v + glo + c0;
                double x[kItems];
#pragma unroll
                for (int j = 0; j < kItems; ++j) {
                    const int q = j * kBlock + static_cast<int>(threadIdx.x);
                    x[j] = vc[q < cn ? q : cn - 1];
                }
#pragma unroll
                for (int j = 0; j < kItems; ++j) {
                    keep_loaded(x[j]);
                    vs[j * kBlock + threadIdx.x] = x[j];
                }
            }
            __syncthreads();
            const double vlast = vs[cn - 1];
            int lo = -1;   // blocked: the previous answer in this chunk
#pragma unroll
            for (int j = 0; j < kItems; ++j) {
                if ((open >> j) & 1u) {
                    const int k = dense ? j * kBlock + static_cast<int>(threadIdx.x) : kb + j;
                    const double sk = tile[lds_pad(k)];
                    if (sk < vlast) {
                        // first index with v > sk in [a, cn); vs[cn - 1] > sk
                        int a = (dense || lo < 0) ? 0 : lo;
                        int steps = (dense || lo < 0) ? 8 : 0;
                        while (steps < 8 && vs[a] <= sk) {
                            ++a;
                            ++steps;
                        }
                        if (steps == 8) {
                            int hi = cn - 1;
                            while (a < hi) {
                                const int mid = (a + hi) >> 1;
                                if (vs[mid] <= sk)
                                    a = mid + 1;
                                else
                                    hi = mid;
                            }
                        }
                        lo = a;
                        cnt[j] = c0 + a;
                        open &= ~(1u << j);
                    }
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const int k = dense ? j * kBlock + static_cast<int>(threadIdx.x) : kb + j;
            if (k < nvalid) {
                const long long g = (base + k == M - 1) ? nd : glo + cnt[j];
                tile[lds_pad(k)] = __builtin_bit_cast(double, g);
            }
        }
        if (threadIdx.x == 0)
            gfirst = glo;
    } else {
        for (int j = 0; j < kItems; ++j) {
            const int k = j * kBlock + static_cast<int>(threadIdx.x);
            if (k < nvalid) {
                const double sk = tile[lds_pad(k)];
                if (k == nvalid - 1)
                    xend = sk;
                const long long g = (base + k == M - 1) ? nd : us.count(sk);
                tile[lds_pad(k)] = __builtin_bit_cast(double, g);
            }
        }
    }
    __syncthreads();
    const auto sv = src.loaded();
    long long z = 0, e = 0, c = 0;
    // the integer parts (residual schemes; zero otherwise) from clamped,
    // unconditional loads, all in flight before their first use
    long long ig[kItems];
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const int k = j * kBlock + static_cast<int>(threadIdx.x);
        ig[j] = sv.integ(base + (k < nvalid ? k : nvalid - 1));
    }
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const int k = j * kBlock + static_cast<int>(threadIdx.x);
        keep_loaded(ig[j]);
        if (k < nvalid) {
            const long long g = __builtin_bit_cast(long long, tile[lds_pad(k)]);
            const long long gp = (k == 0) ? gfirst :
                __builtin_bit_cast(long long, tile[lds_pad(k - 1)]);
            const long long r = g - gp + ig[j];
            rep[base + k] = static_cast<int>(r);
            z += r == 0;
            e += r > 1 ? r - 1 : 0;
            c += r > 1;
        }
    }
    if (tagg) {
        // rep -> index aggregates of this tile (childless slots, extra
        // copies, sources with extras)
        z = block_reduce(z, OpAddL(), red);
        e = block_reduce(e, OpAddL(), red);
        c = block_reduce(c, OpAddL(), red);
        if (threadIdx.x == 0) {
            tagg[3 * tidx + 0] = z;
            tagg[3 * tidx + 1] = e;
            tagg[3 * tidx + 2] = c;
        }
    }
    __syncthreads();
    const double xe = xend;
    __syncthreads();
    return xe;
}

// Writes f(s_i) for the exact prefix s (tests; sorted uniforms).
struct EpiRaw {
    double *out;
    __device__ void operator()(long long i, double s) const { out[i] = s; }
};
// u_j = 1 - exp(lmax_j) with lmax_j = -s_j (u01_sequence.hpp:151)
struct EpiSorted {
    double *out;
    __device__ void operator()(long long i, double s) const
    {
        out[i] = 1.0 - exp(-s);
    }
};

template <class WSrc, class Epi>
__global__ void __launch_bounds__(kBlock) k_prefix(WSrc src, long long nh,
    const long long *ndev, const ResampleDev *prm, FoldBufs fb, Epi epi)
{
    if (!prm->active)
        return;
    const long long n = eff_n(ndev, nh);
    const long long base = static_cast<long long>(blockIdx.x) * kTile;
    if (base >= n)
        return;
    __shared__ double tile[kTileLds];
    __shared__ fold::ExactShared ex;
    load_tile_lds(src, base, n, tile, 0.0);
    fold::tile_exact(tile, fb.xstart[blockIdx.x], ex);
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const int k = j * kBlock + static_cast<int>(threadIdx.x);
        if (base + k < n)
            epi(base + k, tile[lds_pad(k)]);
    }
}

// ------------------------------------------------- running maximum (v)

__global__ void __launch_bounds__(kBlock) k_tile_max(const double *u,
    long long nh, const long long *ndev, const ResampleDev *prm, double *tmax)
{
    if (!prm->active)
        return;
    const long long n = eff_n(ndev, nh);
    const long long base = static_cast<long long>(blockIdx.x) * kTile;
    if (base >= n)
        return;
    __shared__ double lds[kWaves];
    double m = -INFINITY;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const long long g = base + j * kBlock + threadIdx.x;
        if (g < n) {
            const double x = u[g];
            m = x > m ? x : m;
        }
    }
    m = block_reduce(m, OpMaxD(), lds);
    if (threadIdx.x == 0)
        tmax[blockIdx.x] = m;
}

__global__ void __launch_bounds__(kBlock) k_scan_max(const double *tmax,
    long long nh, const long long *ndev, const ResampleDev *prm, double *tpre)
{
    if (!prm->active)
        return;
    const long long n = eff_n(ndev, nh);
    const long long ntiles = ceil_div(n, kTile);
    __shared__ double lds[kWaves];
    double carry = -INFINITY;
    for (long long c = 0; c < ntiles; c += kBlock) {
        const long long t = c + threadIdx.x;
        const double v = t < ntiles ? tmax[t] : -INFINITY;
        double tot;
        const double ex = block_exclusive_scan(v, OpMaxD(), -__builtin_huge_val(), lds, &tot);
        if (t < ntiles)
            tpre[t] = carry > ex ? carry : ex;
        carry = carry > tot ? carry : tot;
    }
}

__global__ void __launch_bounds__(kBlock) k_apply_max(const double *u,
    long long nh, const long long *ndev, const ResampleDev *prm,
    const double *tpre, double *v)
{
    if (!prm->active)
        return;
    const long long n = eff_n(ndev, nh);
    const long long base = static_cast<long long>(blockIdx.x) * kTile;
    if (base >= n)
        return;
    __shared__ double tile[kTileLds];
    __shared__ double lds[kWaves];
    double w[kItems];
    load_tile_blocked(WArray{u}, base, n, w, tile, -INFINITY);
    double m = -INFINITY;
#pragma unroll
    for (int j = 0; j < kItems; ++j)
        m = w[j] > m ? w[j] : m;
    double tot;
    double run = block_exclusive_scan(m, OpMaxD(), -__builtin_huge_val(), lds, &tot);
    const double tp = tpre[blockIdx.x];
    run = run > tp ? run : tp;
    const long long ib = base + static_cast<long long>(threadIdx.x) * kItems;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        run = w[j] > run ? w[j] : run;
        if (ib + j < n)
            v[ib + j] = run;
    }
}

// ------------------------------------------------------ rep -> index

// Per-tile aggregates of rep: (childless slots, extra copies, sources).
// Expand mode (M != N): (unused, copies, sources with copies).
template <bool kExpand>
__global__ void __launch_bounds__(kBlock) k_ridx_agg(const int *rep,
    long long M, const ResampleDev *prm, long long *tagg)
{
    if (prm && !prm->active)
        return;
    const long long base = static_cast<long long>(blockIdx.x) * kTile;
    __shared__ long long lds[kWaves];
    long long z = 0, e = 0, c = 0;
    // all loads first (clamped), then the counts: a load under a branch
    // would wait for itself
    int rv[kItems];
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const long long g = base + j * kBlock + threadIdx.x;
        rv[j] = rep[g < M ? g : M - 1];
    }
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        keep_loaded(rv[j]);
        const long long g = base + j * kBlock + threadIdx.x;
        if (g < M) {
            const long long r = rv[j];
            if (kExpand) {
                e += r;
                c += r > 0;
            } else {
                z += r == 0;
                e += r > 1 ? r - 1 : 0;
                c += r > 1;
            }
        }
    }
    z = block_reduce(z, OpAddL(), lds);
    e = block_reduce(e, OpAddL(), lds);
    c = block_reduce(c, OpAddL(), lds);
    if (threadIdx.x == 0) {
        tagg[3 * blockIdx.x + 0] = z;
        tagg[3 * blockIdx.x + 1] = e;
        tagg[3 * blockIdx.x + 2] = c;
    }
}

struct L3 {
    long long a, b, c;
};
struct I3 {
    int a, b, c;
};
struct OpAddI3 {
    __device__ I3 operator()(I3 x, I3 y) const
    {
        return I3{x.a + y.a, x.b + y.b, x.c + y.c};
    }
};
struct OpAddL3 {
    __device__ L3 operator()(L3 x, L3 y) const
    {
        return L3{x.a + y.a, x.b + y.b, x.c + y.c};
    }
};

template <int kThreads, int kPer>
__device__ __forceinline__ void scan_tiles_block(const long long *tagg, long long ntiles,
    long long *toff);

// One block of kScanThreads: the scan is latency-bound, so four times the
// threads of a tile block, four tiles each.
constexpr int kScanThreads = 1024;

__global__ void __launch_bounds__(kScanThreads) k_ridx_scan(const long long *tagg,
    long long ntiles, const ResampleDev *prm, long long *toff)
{
    if (prm && !prm->active)
        return;
    scan_tiles_block<kScanThreads, 4>(tagg, ntiles, toff);
}

// Exclusive offsets of the per-tile (childless, extras, sources) triples,
// totals at toff[3 ntiles ..]. One block of kThreads, kPer tiles a thread
// per chunk.
template <int kThreads, int kPer>
__device__ __forceinline__ void scan_tiles_block(const long long *tagg, long long ntiles,
    long long *toff)
{
    __shared__ L3 lds[kThreads / kWave];
    const L3 carry = block_scan_chunked_n<kThreads, kPer>(ntiles,
        [&](long long t) { return L3{tagg[3 * t], tagg[3 * t + 1], tagg[3 * t + 2]}; },
        [&](long long t, L3 ex) {
            toff[3 * t + 0] = ex.a;
            toff[3 * t + 1] = ex.b;
            toff[3 * t + 2] = ex.c;
        },
        OpAddL3(), L3{0, 0, 0}, lds);
    if (threadIdx.x == 0) {
        toff[3 * ntiles + 0] = carry.a;
        toff[3 * ntiles + 1] = carry.b;
        toff[3 * ntiles + 2] = carry.c;
    }
}

template <bool kExpand>
__device__ __forceinline__ void scatter_tile(const int *rep, long long M, long long tidx,
    const long long *toff, int *zdst, int *cE, int *csrc, int *idx);

template <bool kExpand>
__global__ void __launch_bounds__(kBlock) k_ridx_scatter(const int *rep,
    long long M, const ResampleDev *prm, const long long *toff, int *zdst,
    int *cE, int *csrc, int *idx)
{
    if (prm && !prm->active)
        return;
    scatter_tile<kExpand>(rep, M, blockIdx.x, toff, zdst, cE, csrc, idx);
}

// The childless slots (zdst, in slot order) and the sources with extra
// copies (csrc, with cE the exclusive prefix of their extra copies) of tile
// tidx of rep, at the offsets toff[3 tidx ..] of the tile.
template <bool kExpand>
__device__ __forceinline__ void scatter_tile(const int *rep, long long M, long long tidx,
    const long long *toff, int *zdst, int *cE, int *csrc, int *idx)
{
    const long long base = tidx * kTile;
    __shared__ L3 lds[kWaves];
    __shared__ int tile[kTileLds];   // rep, element k at lds_pad(k)
    __shared__ int stage[kTile];     // one output list at a time
    const int nvalid = M - base < kTile ? static_cast<int>(M - base) : kTile;
    // all loads first (clamped), then their uses: a load under a branch
    // would wait for itself
    int rl[kItems];
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const int k = j * kBlock + static_cast<int>(threadIdx.x);
        rl[j] = rep[base + (k < nvalid ? k : nvalid - 1)];
    }
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const int k = j * kBlock + static_cast<int>(threadIdx.x);
        keep_loaded(rl[j]);
        const int rv = k < nvalid ? rl[j] : 0;
        tile[lds_pad(k)] = rv;
        if (!kExpand && idx && k < nvalid && rv > 0)
            idx[base + k] = static_cast<int>(base + k);
    }
    __syncthreads();
    // blocked order: thread t owns elements t*kItems ..
    const int kb = static_cast<int>(threadIdx.x) * kItems;
    int r[kItems];
    L3 loc{0, 0, 0};
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const bool valid = kb + j < nvalid;
        r[j] = valid ? tile[lds_pad(kb + j)] : 0;
        const long long rr = r[j];
        if (kExpand) {
            loc.b += rr;
            loc.c += rr > 0;
        } else {
            loc.a += valid && rr == 0;
            loc.b += rr > 1 ? rr - 1 : 0;
            loc.c += rr > 1;
        }
    }
    L3 tot;
    const L3 ex = block_exclusive_scan(loc, OpAddL3(), L3{0, 0, 0}, lds, &tot);
    // each list is staged in LDS in tile order and written out coalesced
    if (!kExpand && zdst) {
        int zr = static_cast<int>(ex.a);
#pragma unroll
        for (int j = 0; j < kItems; ++j)
            if (kb + j < nvalid && r[j] == 0)
                stage[zr++] = static_cast<int>(base + kb + j);
        __syncthreads();
        const long long z0 = toff[3 * tidx + 0];
        for (int q = threadIdx.x; q < tot.a; q += kBlock)
            zdst[z0 + q] = stage[q];
        __syncthreads();
    }
    const long long c0 = toff[3 * tidx + 2];
    {
        int cr = static_cast<int>(ex.c);
        long long er = toff[3 * tidx + 1] + ex.b;
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const int rr = r[j];
            if (kExpand ? rr > 0 : rr > 1) {
                stage[cr++] = static_cast<int>(er);
                er += kExpand ? rr : rr - 1;
            }
        }
        __syncthreads();
        for (int q = threadIdx.x; q < tot.c; q += kBlock)
            cE[c0 + q] = stage[q];
        __syncthreads();
    }
    {
        int cr = static_cast<int>(ex.c);
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const int rr = r[j];
            if (kExpand ? rr > 0 : rr > 1)
                stage[cr++] = static_cast<int>(base + kb + j);
        }
        __syncthreads();
        for (int q = threadIdx.x; q < tot.c; q += kBlock)
            csrc[c0 + q] = stage[q];
        __syncthreads();
    }
}

__device__ inline long long upper_bound_i(const int *a, long long n, long long x)
{
    long long lo = 0, hi = n;
    while (lo < hi) {
        const long long mid = (lo + hi) >> 1;
        if (a[mid] <= x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// The same search by one whole wave, 64 probes a round: about log64(n)
// dependent loads instead of log2(n) (every lane returns the result; the
// loop bounds are wave-uniform).
__device__ inline long long wave_upper_bound_i(const int *a, long long n, long long x)
{
    const long long lane = threadIdx.x & (kWave - 1);
    long long lo = 0, hi = n;   // the first index with a[i] > x lies in [lo, hi]
    while (lo < hi) {
        const long long step = (hi - lo + kWave - 1) / kWave;
        const long long p = lo + lane * step;
        const bool le = p < hi && a[p] <= x;
        // a is non-decreasing, so the probes with a[p] <= x are a prefix
        const long long c = __popcll(__ballot(le));
        if (c == 0)
            break;                                  // a[lo] > x
        const long long nhi = lo + c * step;
        lo = lo + (c - 1) * step + 1;               // past the last such probe
        hi = nhi < hi ? nhi : hi;                   // at the next probe, if any
    }
    return lo;
}

// Source of each rank k (the k-th childless slot, or the k-th output in
// expand mode): the source whose copies cover k. Optionally gathers an SoA
// state in place: x[d][dst] = x[d][src] (sources are never destinations).
template <bool kExpand, typename E>
__device__ __forceinline__ void expand_chunk(long long chunk, long long K, long long C,
    const int *zdst, const int *cE, const int *csrc, int *idx, E *state,
    int dim, long long ld);

// With end.pd, block 0 also applies the event's end to the particle set
// (after_event on eprm): nothing in this kernel reads what it writes.
template <bool kExpand, typename E>
__global__ void __launch_bounds__(kBlock) k_expand(const ResampleDev *prm,
    const long long *toff, long long ntiles, const int *zdst, const int *cE,
    const int *csrc, int *idx, E *state, int dim, long long ld,
    ResampleDev *eprm, EventEnd end)
{
    if (prm && !prm->active)
        return;
    if (end.pd && blockIdx.x == 0 && threadIdx.x == 0)
        after_event(end.pd, eprm, end.n, end.rep, end.scheme, end.words,
            end.deferred);
    const long long K = kExpand ? toff[3 * ntiles + 1] : toff[3 * ntiles + 0];
    const long long C = toff[3 * ntiles + 2];
    if (static_cast<long long>(blockIdx.x) * kTile >= K || C <= 0)
        return;
    expand_chunk<kExpand, E>(blockIdx.x, K, C, zdst, cE, csrc, idx, state, dim,
        ld);
}

// Ranks [chunk kTile, (chunk + 1) kTile) of the K ranks: the source of each
// rank is the one of the C sources whose copies cover it.
template <bool kExpand, typename E>
__device__ __forceinline__ void expand_chunk(long long chunk, long long K, long long C,
    const int *zdst, const int *cE, const int *csrc, int *idx, E *state,
    int dim, long long ld)
{
    const long long k0 = chunk * kTile;
    const int nk = static_cast<int>((k0 + kTile < K ? k0 + kTile : K) - k0);
    __shared__ int mark[kTileLds];
    __shared__ int mred[kWaves];
    __shared__ long long clo_s, chi_s;
    // the first and last source of the chunk, one wave each
    const int wave = static_cast<int>(threadIdx.x) / kWave;
    if (wave < 2) {
        const long long b = wave_upper_bound_i(cE, C, wave == 0 ? k0 : k0 + nk - 1) - 1;
        if ((threadIdx.x & (kWave - 1)) == 0) {
            if (wave == 0)
                clo_s = b < 0 ? 0 : b;
            else
                chi_s = b < 0 ? 0 : b;
        }
    }
#pragma unroll
    for (int j = 0; j < kItems; ++j)
        mark[lds_pad(j * kBlock + static_cast<int>(threadIdx.x))] = -1;
    __syncthreads();
    const long long clo = clo_s, chi = chi_s;
    // the sources covering this chunk's ranks start at strictly increasing
    // ranks (each covers at least one), so there are at most kTile of them;
    // otherwise (counts that break this) search per rank
    const bool marked = chi - clo + 1 <= kTile;
    if (marked) {
        // each source marks the first rank of the chunk it covers ...
        for (long long q = clo + threadIdx.x; q <= chi; q += kBlock) {
            long long p = static_cast<long long>(cE[q]) - k0;
            p = p < 0 ? 0 : p;
            if (p < nk)
                mark[lds_pad(static_cast<int>(p))] = static_cast<int>(q - clo);
        }
        __syncthreads();
        // ... and a running maximum carries it over the ranks it covers
        const int kb = static_cast<int>(threadIdx.x) * kItems;
        int run = -1;
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const int v = mark[lds_pad(kb + j)];
            run = v > run ? v : run;
        }
        int tot;
        run = block_exclusive_scan(run, OpMaxI(), -1, mred, &tot);
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const int v = mark[lds_pad(kb + j)];
            run = v > run ? v : run;
            mark[lds_pad(kb + j)] = run;
        }
        __syncthreads();
    }
    // all index loads of this thread's ranks first, then all state loads,
    // then the stores: the memory latencies overlap instead of adding up
    // (loads from clamped ranks, unconditional: a guarded load would wait
    // for itself)
    // src[j] is a valid source for every j (an idle rank takes the last
    // rank's); bit j of live marks the ranks below nk
    int src[kItems], dst[kItems];
    unsigned live = 0;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const int kr = j * kBlock + static_cast<int>(threadIdx.x);
        const int kk = kr < nk ? kr : nk - 1;
        const long long k = k0 + kk;
        long long c = marked ? clo + mark[lds_pad(kk)] : upper_bound_i(cE, C, k) - 1;
        c = c < 0 ? 0 : (c >= C ? C - 1 : c);
        src[j] = csrc[c];
        dst[j] = kExpand ? static_cast<int>(k) : zdst[k];
        live |= kr < nk ? 1u << j : 0u;
    }
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        keep_loaded(src[j]);
        keep_loaded(dst[j]);
    }
    if (idx) {
#pragma unroll
        for (int j = 0; j < kItems; ++j)
            if ((live >> j) & 1u)
                idx[dst[j]] = src[j];
    }
    if (state) {
        // sources are never destinations, so the loads see the old state;
        // one component at a time, from a uniform column base
        for (int d = 0; d < dim; ++d) {
            E *const col = state + d * ld;
            E v[kItems];
            // unconditional loads (every src[j] is valid)
#pragma unroll
            for (int j = 0; j < kItems; ++j)
                v[j] = col[static_cast<unsigned>(src[j])];
#pragma unroll
            for (int j = 0; j < kItems; ++j)
                keep_loaded(v[j]);
#pragma unroll
            for (int j = 0; j < kItems; ++j)
                if ((live >> j) & 1u)
                    col[static_cast<unsigned>(dst[j])] = v[j];
        }
    }
    __syncthreads();
}

// Tiles whose childless, extras and sources prefixes k_ridx_prefix holds in
// LDS (three int arrays, 60 KB) to find every tile's slice of sources there.
constexpr int kPrefixLdsTiles = 5120;

// Last index i in [0, n) with a[i] <= x for a non-decreasing a in LDS, or
// -1 (one thread).
__device__ inline int lds_last_le(const int *a, int n, long long x)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] <= x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo - 1;
}

// The M == N rep -> index prefixes, one block: the exclusive per-tile
// offsets of (childless, extras, sources) as k_ridx_scan (toff, totals at
// [3 ntiles ..]) and, when the tiles fit kPrefixLdsTiles, for every tile t
// with childless slots the slice [tq[2t], tq[2t + 1]) of the compacted
// sources they take from: from the first source of the tile whose extras
// hold rank Z_t to the last source of the tile whose extras hold rank
// Z_{t+1} - 1, searched in LDS. Per-tile counts and their prefixes stay
// below 2^31 (M < 2^31).
__global__ void __launch_bounds__(kScanThreads) k_ridx_prefix(const long long *tagg,
    long long ntiles, const ResampleDev *prm, long long *toff, int *tq)
{
    if (prm && !prm->active)
        return;
    constexpr int kPer = 4;   // tiles per thread and chunk
    __shared__ I3 red[kScanThreads / kWave];
    __shared__ int zp[kPrefixLdsTiles + 1];   // childless prefix
    __shared__ int ep[kPrefixLdsTiles + 1];   // extras prefix
    __shared__ int cp[kPrefixLdsTiles + 1];   // sources prefix
    const bool in_lds = tq != nullptr && ntiles <= kPrefixLdsTiles;
    I3 carry{0, 0, 0};
    for (long long c0 = 0; c0 < ntiles; c0 += static_cast<long long>(kScanThreads) * kPer) {
        const long long b0 = c0 + static_cast<long long>(threadIdx.x) * kPer;
        // clamped, unconditional loads, all in flight before their first use
        long long raw[3 * kPer];
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const long long t = b0 + k < ntiles ? b0 + k : ntiles - 1;
            raw[3 * k + 0] = tagg[3 * t + 0];
            raw[3 * k + 1] = tagg[3 * t + 1];
            raw[3 * k + 2] = tagg[3 * t + 2];
        }
        I3 v[kPer];
        I3 loc{0, 0, 0};
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            keep_loaded(raw[3 * k + 0]);
            keep_loaded(raw[3 * k + 1]);
            keep_loaded(raw[3 * k + 2]);
            const bool ok = b0 + k < ntiles;
            v[k] = ok ? I3{static_cast<int>(raw[3 * k]), static_cast<int>(raw[3 * k + 1]),
                static_cast<int>(raw[3 * k + 2])} : I3{0, 0, 0};
            loc = OpAddI3()(loc, v[k]);
        }
        I3 tot;
        I3 run = OpAddI3()(carry,
            block_exclusive_scan_n<kScanThreads>(loc, OpAddI3(), I3{0, 0, 0}, red, &tot));
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const long long t = b0 + k;
            if (t < ntiles) {
                toff[3 * t + 0] = run.a;
                toff[3 * t + 1] = run.b;
                toff[3 * t + 2] = run.c;
                if (in_lds) {
                    zp[t] = run.a;
                    ep[t] = run.b;
                    cp[t] = run.c;
                }
            }
            run = OpAddI3()(run, v[k]);
        }
        carry = OpAddI3()(carry, tot);
    }
    if (threadIdx.x == 0) {
        toff[3 * ntiles + 0] = carry.a;
        toff[3 * ntiles + 1] = carry.b;
        toff[3 * ntiles + 2] = carry.c;
        if (in_lds) {
            zp[ntiles] = carry.a;
            ep[ntiles] = carry.b;
            cp[ntiles] = carry.c;
        }
    }
    if (!in_lds)
        return;
    __syncthreads();   // zp, ep and cp complete for the whole block
    const int nt = static_cast<int>(ntiles);
    for (int t = threadIdx.x; t < nt; t += kScanThreads) {
        const long long Z = zp[t];
        const long long Z1 = zp[t + 1];
        const bool anyz = Z1 > Z;
        int slo = lds_last_le(ep, nt, Z);
        int shi = lds_last_le(ep, nt, anyz ? Z1 - 1 : Z);
        slo = slo < 0 ? 0 : slo;
        shi = shi < 0 ? 0 : shi;
        tq[2 * t] = anyz ? cp[slo] : 0;
        tq[2 * t + 1] = anyz ? cp[shi + 1] : 0;
    }
}

// The sources with extra copies of tile tidx, in order (csrc), and the
// exclusive prefix of their extra copies (cE), at the tile's offsets in
// toff (M == N). The offsets load with the tile's rep; one block scan of
// the packed per-thread (extras << 16 | sources).
__global__ void __launch_bounds__(kBlock) k_ridx_sources(const int *rep,
    long long M, const ResampleDev *prm, const long long *toff, int *cE,
    int *csrc)
{
    if (prm && !prm->active)
        return;
    const long long tidx = blockIdx.x;
    const long long base = tidx * kTile;
    const int nvalid = M - base < kTile ? static_cast<int>(M - base) : kTile;
    __shared__ int lds[kTileLds];
    __shared__ long long red[kWaves];
    const long long e0 = toff[3 * tidx + 1];
    const long long c0 = toff[3 * tidx + 2];
    const int *const rept = rep + base;
    {
        // clamped, unconditional loads, all in flight before their first use
        int v[kItems];
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const int k = j * kBlock + static_cast<int>(threadIdx.x);
            v[j] = rept[static_cast<unsigned>(k < nvalid ? k : nvalid - 1)];
        }
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const int k = j * kBlock + static_cast<int>(threadIdx.x);
            keep_loaded(v[j]);
            lds[lds_pad(k)] = k < nvalid ? v[j] : 0;
        }
    }
    __syncthreads();
    const int kb = static_cast<int>(threadIdx.x) * kItems;
    int r[kItems];
    long long ec = 0;   // extras << 16 | sources (at most kItems sources)
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        r[j] = lds[lds_pad(kb + j)];
        ec += r[j] > 1 ? (static_cast<long long>(r[j] - 1) << 16) | 1 : 0;
    }
    long long tot;
    const long long ex = block_exclusive_scan(ec, OpAddL(), 0LL, red, &tot);
    // (the scan ends with a barrier: every rep has been read)
    const int nsrc = static_cast<int>(tot & 0xffff);   // sources of the tile
    // cE, staged in tile order, then written out coalesced
    {
        int cr = static_cast<int>(ex & 0xffff);
        long long er = e0 + (ex >> 16);
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            if (r[j] > 1) {
                lds[cr++] = static_cast<int>(er);
                er += r[j] - 1;
            }
        }
        __syncthreads();
        for (int q = threadIdx.x; q < nsrc; q += kBlock)
            cE[c0 + q] = lds[q];
        __syncthreads();
    }
    {
        int cr = static_cast<int>(ex & 0xffff);
#pragma unroll
        for (int j = 0; j < kItems; ++j)
            if (r[j] > 1)
                lds[cr++] = static_cast<int>(base + kb + j);
        __syncthreads();
        for (int q = threadIdx.x; q < nsrc; q += kBlock)
            csrc[c0 + q] = lds[q];
    }
}

// Ancestor indices of tile tidx when M == N (the rep -> index layout of
// transform.hpp:116-133): a particle with copies keeps its own index; the
// tile's childless slots, ranks [Z, Z + z) among all childless slots in
// order, take the sources whose extra copies cover those ranks (csrc, with
// cE the exclusive prefix of their extra copies). Those sources lie in the
// compacted slices of the source tiles whose extras cover the ranks: the
// slice comes from tq (k_ridx_prefix) or, without it, from the per-tile
// prefixes (toff) in two rounds of a wave search. The block scans that
// slice coalesced, and the tile's indices are written in one coalesced pass.
// With `state`, the childless slots' rows are gathered in place (sources
// always have copies, so they are never written). Block 0 ends the event
// when end.pd is set; nothing in this kernel reads what that writes.
// kReg: the sources come per tile (csrc / cE hold tile u's sources and
// their extras prefix within the tile at u kTile .., written by the fast
// event's counting pass) instead of compacted over all tiles.
template <typename E, bool kReg = false>
__global__ void __launch_bounds__(kBlock, 5) k_ridx_fill(const ResampleDev *prm,
    const int *rep, long long M, const long long *toff, const int *tq,
    long long ntiles, const int *cE, const int *csrc, int *idx, E *state,
    int dim, long long ld, ResampleDev *eprm, EventEnd end)
{
    if (prm && !prm->active)
        return;
    if (end.pd && blockIdx.x == 0 && threadIdx.x == 0)
        after_event(end.pd, eprm, end.n, end.rep, end.scheme, end.words,
            end.deferred);
    const long long tidx = blockIdx.x;
    const long long base = tidx * kTile;
    const int nvalid = M - base < kTile ? static_cast<int>(M - base) : kTile;
    // one array, reused: rep, then per childless rank the slot of its source
    // in csrc, then the output index
    __shared__ int lds[kTileLds];
    __shared__ int ired[kWaves];
    __shared__ long long qrange[2];
    __shared__ int src0s;   // the source of the tile's first childless rank
    // the tile's childless ranks [Z, Z + zt) and the sources' count C come
    // from the per-tile prefixes, so the searches for the source tiles start
    // at once, with the tile's rep loads in flight
    const long long Z = toff[3 * tidx + 0];
    const long long zt = toff[3 * (tidx + 1) + 0] - Z;
    const long long C = toff[3 * ntiles + 2];
    const bool any = zt > 0 && C > 0;
    // uniform bases and 32-bit offsets: one VGPR per address
    const int *const rept = rep + base;
    // clamped, unconditional loads, all in flight before their first use
    // (a guarded load waits for itself); padding counts as kept
    {
        int v[kItems];
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const int k = j * kBlock + static_cast<int>(threadIdx.x);
            v[j] = rept[static_cast<unsigned>(k < nvalid ? k : nvalid - 1)];
        }
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const int k = j * kBlock + static_cast<int>(threadIdx.x);
            keep_loaded(v[j]);
            lds[lds_pad(k)] = k < nvalid ? v[j] : 1;
        }
    }
    const int wave = static_cast<int>(threadIdx.x) / kWave;
    if (!kReg && tq) {
        if (threadIdx.x < 2)
            qrange[threadIdx.x] = tq[2 * tidx + threadIdx.x];
    } else if (any && wave < 2) {
        // the source tiles whose extras hold the first and the last
        // childless rank: the last tile whose extras start at or before it
        const long long x = wave == 0 ? Z : Z + zt - 1;
        long long sx = wave_last_le([&](long long t) { return toff[3 * t + 1]; },
            ntiles, x);
        sx = sx < 0 ? 0 : sx;
        if ((threadIdx.x & (kWave - 1)) == 0) {
            if (kReg)
                qrange[wave] = sx;
            else
                qrange[wave] = wave == 0 ? toff[3 * sx + 2] : toff[3 * (sx + 1) + 2];
        }
    }
    __syncthreads();
    // blocked order: thread t owns elements t kItems .. t kItems + kItems - 1;
    // bit j of cmask: element kb + j is childless
    const int kb = static_cast<int>(threadIdx.x) * kItems;
    unsigned cmask = 0;
    int z = 0;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const bool c = lds[lds_pad(kb + j)] == 0;
        cmask |= c ? 1u << j : 0u;
        z += c;
    }
    int ztot;
    const int zex = block_exclusive_scan(z, fold::OpAddI(), 0, ired, &ztot);
    // (the scan ends with a barrier: every rep has been read; ztot == zt
    // for counts that sum to M)
#pragma unroll
    for (int j = 0; j < kItems; ++j)
        lds[lds_pad(j * kBlock + static_cast<int>(threadIdx.x))] = -1;
    __syncthreads();
    int out[kItems];
    if (any && ztot > 0) {
        const long long q0 = qrange[0];
        const long long q1 = qrange[1] < C ? qrange[1] : C;
        // each source starting after Z marks its first rank of the tile; the
        // one covering rank Z (the last starting at or before it) marks rank
        // 0, by a maximum. All loads of a thread in flight at once, a tile's
        // worth of sources per pass.
        int first = -1;
        if (kReg) {
            // source tiles q0 .. qrange[1]; each rank holds its source's
            // index itself
            for (long long u = q0; u <= qrange[1]; ++u) {
                const int *const cEu = cE + u * kTile;
                const int *const csu = csrc + u * kTile;
                const long long Eu = toff[3 * u + 1];
                const int nq = static_cast<int>(toff[3 * (u + 1) + 2] - toff[3 * u + 2]);
                if (nq <= 0)
                    continue;
                int e[kItems], sid[kItems];
#pragma unroll
                for (int j = 0; j < kItems; ++j) {
                    const int o = j * kBlock + static_cast<int>(threadIdx.x);
                    const unsigned oc = static_cast<unsigned>(o < nq ? o : nq - 1);
                    e[j] = cEu[oc];
                    sid[j] = csu[oc];
                }
#pragma unroll
                for (int j = 0; j < kItems; ++j) {
                    const int o = j * kBlock + static_cast<int>(threadIdx.x);
                    keep_loaded(e[j]);
                    keep_loaded(sid[j]);
                    if (o < nq) {
                        const long long pq = Eu + e[j] - Z;
                        if (pq <= 0)
                            first = sid[j] > first ? sid[j] : first;
                        else if (pq < ztot)
                            lds[lds_pad(static_cast<int>(pq))] = sid[j];
                    }
                }
            }
        }
        for (long long c0 = q0; c0 < (kReg ? q0 : q1); c0 += kTile) {
            const int *const cEc = cE + c0;
            const int nq = q1 - c0 < kTile ? static_cast<int>(q1 - c0) : kTile;
            int e[kItems];
#pragma unroll
            for (int j = 0; j < kItems; ++j) {
                const int o = j * kBlock + static_cast<int>(threadIdx.x);
                e[j] = cEc[static_cast<unsigned>(o < nq ? o : nq - 1)];
            }
#pragma unroll
            for (int j = 0; j < kItems; ++j) {
                const int o = j * kBlock + static_cast<int>(threadIdx.x);
                keep_loaded(e[j]);
                e[j] = o < nq ? e[j] : INT_MAX;
            }
#pragma unroll
            for (int j = 0; j < kItems; ++j) {
                const int q = static_cast<int>(c0) + j * kBlock + static_cast<int>(threadIdx.x);
                const long long pq = static_cast<long long>(e[j]) - Z;
                if (pq <= 0)
                    first = q > first ? q : first;
                else if (pq < ztot)
                    lds[lds_pad(static_cast<int>(pq))] = q;
            }
        }
        first = wave_reduce(first, OpMaxI());
        if ((threadIdx.x & (kWave - 1)) == 0 && first >= 0)
            atomicMax(&lds[lds_pad(0)], first);
        __syncthreads();
        // ... and a running maximum carries them over the ranks they cover
        int run = -1;
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const int v = lds[lds_pad(kb + j)];
            run = v > run ? v : run;
        }
        int tot;
        run = block_exclusive_scan(run, OpMaxI(), -1, ired, &tot);
        if (threadIdx.x == 0) {
            int q = lds[lds_pad(0)];
            if (kReg)
                src0s = q < 0 ? static_cast<int>(base) : q;
            else {
                q = q < 0 ? 0 : (q >= C ? static_cast<int>(C - 1) : q);
                src0s = csrc[static_cast<unsigned>(q)];
            }
        }
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const int v = lds[lds_pad(kb + j)];
            run = v > run ? v : run;
            lds[lds_pad(kb + j)] = run;
        }
        __syncthreads();
        // every element loads (a kept one the source of a valid rank), so
        // that no load sits behind a branch
        int zr = zex;
        int sv[kItems];
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const bool c = (cmask >> j) & 1u;
            int q = lds[lds_pad(zr < ztot ? zr : ztot - 1)];
            if (kReg) {
                // (a rank always has a source for counts that sum to M)
                sv[j] = q < 0 ? static_cast<int>(base + kb + j) : q;
            } else {
                q = q < 0 ? 0 : (q >= C ? static_cast<int>(C - 1) : q);
                sv[j] = csrc[static_cast<unsigned>(q)];
            }
            zr += c;
        }
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            keep_loaded(sv[j]);
            out[j] = (cmask >> j) & 1u ? sv[j] : static_cast<int>(base + kb + j);
        }
        __syncthreads();   // every rank's source read before the outputs go in
    } else {
        // no childless slot (or no source to take one from): identity
#pragma unroll
        for (int j = 0; j < kItems; ++j)
            out[j] = static_cast<int>(base + kb + j);
    }
#pragma unroll
    for (int j = 0; j < kItems; ++j)
        lds[lds_pad(kb + j)] = out[j];
    __syncthreads();
    int *const idxt = idx ? idx + base : nullptr;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const int k = j * kBlock + static_cast<int>(threadIdx.x);
        out[j] = lds[lds_pad(k)];
        if (idxt && k < nvalid)
            idxt[static_cast<unsigned>(k)] = out[j];
    }
    if (state && any && ztot > 0) {
        const int src0 = src0s;
        // one component at a time: all loads of a thread, then its stores
        for (int d = 0; d < dim; ++d) {
            E *const col = state + d * ld;
            E *const colt = col + base;
            E v[kItems];
            // unconditional loads: a kept or padding element reads the
            // tile's first source (one line, shared) instead of waiting
#pragma unroll
            for (int j = 0; j < kItems; ++j) {
                const int k = j * kBlock + static_cast<int>(threadIdx.x);
                const bool c = k < nvalid && out[j] != base + k;
                v[j] = col[static_cast<unsigned>(c ? out[j] : src0)];
            }
#pragma unroll
            for (int j = 0; j < kItems; ++j)
                keep_loaded(v[j]);
#pragma unroll
            for (int j = 0; j < kItems; ++j) {
                const int k = j * kBlock + static_cast<int>(threadIdx.x);
                if (k < nvalid && out[j] != base + k)
                    colt[static_cast<unsigned>(k)] = v[j];
            }
        }
    }
}

__global__ void k_index_rep(const int *idx, long long N, int *rep)
{
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < N)
        atomicAdd(&rep[idx[i]], 1);
}

template <typename T>
__global__ void k_fill(T *p, long long n, T v, const ResampleDev *prm)
{
    if (prm && !prm->active)
        return;
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

// ------------------------------------------------- batched filters

// Batched resampling, the counting stages. Every tile of every resampling
// filter (tile_filter: the filter of each tile) gets its summary from the
// approximate prefix tpre (as k_segments); one workgroup per filter walks
// its tiles' summaries from the exact start 0, folding a tile exactly where
// a summary does not apply (as k_resolve, without groups: a filter has at
// most a few hundred tiles); then every tile counts from its exact start
// with the full exact fold, and one workgroup per filter scans the rep ->
// index aggregates. Same counts as the single-filter pipeline. Filters
// whose prm->active is 0 return at once.
__global__ void __launch_bounds__(kBlock) k_batch_segments(
    const BatchFilter *bf, const int *tile_filter, const ResampleDev *prm,
    const double *l, const double *tpre, TileSeg *segs)
{
    const int f = tile_filter[blockIdx.x];
    if (!prm[f].active)
        return;
    const BatchFilter F = bf[f];
    const long long k = static_cast<long long>(blockIdx.x) - F.tile0;
    const long long base = k * kTile;
    __shared__ fold::SegShared sh;
    load_tile_lds(WLog{l + F.off, prm + f}, base, F.n, sh.tile, 0.0);
    const long long rem = F.n - base;
    const int nvalid = rem < kTile ? static_cast<int>(rem) : kTile;
    fold::tile_segments(nvalid, k == 0 ? 0.0 : tpre[blockIdx.x], &segs[blockIdx.x],
        sh);
}

// Tile summaries a walk holds in LDS at a time.
constexpr int kWalkWin = 16;

__global__ void __launch_bounds__(kBlock) k_batch_walk(const BatchFilter *bf,
    const ResampleDev *prm, const double *l, const TileSeg *segs,
    double *xstart)
{
    const int f = blockIdx.x;
    const ResampleDev *p = prm + f;
    if (!p->active)
        return;
    const BatchFilter F = bf[f];
    const WLog src{l + F.off, p};
    __shared__ double tile[kTileLds];
    __shared__ fold::ExactShared ex;
    __shared__ TileSeg win[kWalkWin];
    __shared__ double xs;
    __shared__ long long next;   // first tile not walked yet
    __shared__ int ok;
    constexpr int kWords = static_cast<int>(sizeof(TileSeg) / 4);
    static_assert(sizeof(TileSeg) % 4 == 0, "summaries copy as words");
    if (threadIdx.x == 0) {
        xs = 0.0;
        next = 0;
    }
    __syncthreads();
    for (long long w0 = 0; w0 < F.ntiles; w0 += kWalkWin) {
        // stage the window's summaries (all threads, coalesced words)
        const int nw = F.ntiles - w0 < kWalkWin ? static_cast<int>(F.ntiles - w0) : kWalkWin;
        const int *from = reinterpret_cast<const int *>(segs + F.tile0 + w0);
        int *to = reinterpret_cast<int *>(win);
        for (int q = threadIdx.x; q < nw * kWords; q += kBlock)
            to[q] = from[q];
        __syncthreads();
        for (;;) {
            // every thread takes its copy of `next` before thread 0 may
            // write it again, so all waves leave the loop together
            const long long cur = next;
            __syncthreads();
            if (cur >= w0 + nw)
                break;
            // thread 0 applies summaries until one does not hold
            if (threadIdx.x == 0) {
                double x = xs;
                long long k = next;
                int good = 1;
                for (; k < w0 + nw; ++k) {
                    xstart[F.tile0 + k] = x;
                    double y = x;
                    if (!fold::apply_segments(win[k - w0], y)) {
                        good = 0;
                        break;
                    }
                    x = y;
                }
                xs = x;
                next = k;
                ok = good;
            }
            __syncthreads();
            if (ok)
                break;
            // tile `next` folded exactly from its exact start xs
            const long long k = next;
            const long long base = k * kTile;
            load_tile_lds(src, base, F.n, tile, 0.0);
            fold::tile_exact(tile, xs, ex);
            const long long rem = F.n - base;
            const int last = (rem < kTile ? static_cast<int>(rem) : kTile) - 1;
            const double y = tile[lds_pad(last)];
            __syncthreads();
            if (threadIdx.x == 0) {
                xs = y;
                next = k + 1;
            }
            __syncthreads();
        }
        __syncthreads();
    }
}

// Counts of every tile of the resampling filters from their exact starts:
// the one-binade fast path, a tile that needs more appended (as its batch
// tile index) to `hard` for k_batch_count_hard.
template <class USrc>
__global__ void __launch_bounds__(kBlock) k_batch_count(const BatchFilter *bf,
    const int *tile_filter, const ResampleDev *prm, const double *l,
    const double *xstart, int *rep, long long *tagg, int *hard)
{
    const int f = tile_filter[blockIdx.x];
    const ResampleDev *p = prm + f;
    if (!p->active)
        return;
    const BatchFilter F = bf[f];
    const long long k = static_cast<long long>(blockIdx.x) - F.tile0;
    count_tile<false>(WLog{l + F.off, p}, USrc{p}, F.n, k, xstart[blockIdx.x],
        rep + F.off, tagg + 3 * F.tile0, hard, static_cast<int>(blockIdx.x));
}

// The listed tiles (hard[1 .. hard[0]]) with the full exact fold.
template <class USrc>
__global__ void __launch_bounds__(kBlock) k_batch_count_hard(
    const BatchFilter *bf, const int *tile_filter, const ResampleDev *prm,
    const double *l, const double *xstart, int *rep, long long *tagg,
    const int *hard)
{
    const int nhard = hard[0];
    for (int q = blockIdx.x; q < nhard; q += gridDim.x) {
        const int id = hard[1 + q];
        const int f = tile_filter[id];
        const ResampleDev *p = prm + f;
        const BatchFilter F = bf[f];
        const long long k = static_cast<long long>(id) - F.tile0;
        count_tile<true>(WLog{l + F.off, p}, USrc{p}, F.n, k, xstart[id],
            rep + F.off, tagg + 3 * F.tile0, static_cast<int *>(nullptr));
    }
}

__global__ void __launch_bounds__(kBlock) k_batch_scan(const BatchFilter *bf,
    const ResampleDev *prm, const long long *tagg, long long *toff)
{
    const int f = blockIdx.x;
    if (!prm[f].active)
        return;
    const BatchFilter F = bf[f];
    scan_tiles_block<kBlock, kItems>(tagg + 3 * F.tile0, F.ntiles,
        toff + 3 * (F.tile0 + f));
}

// The scatter of childless slots and extra copies, one workgroup per tile
// of every filter (tile_filter: the filter of each tile), and the in-place
// gather, one workgroup per chunk of kTile ranks (at most a filter's tile
// count): the per-filter stages that do not chain from tile to tile run
// over the whole device instead of inside the filter's workgroup.
__global__ void __launch_bounds__(kBlock) k_batch_scatter(
    const BatchFilter *bf, const int *tile_filter, const ResampleDev *prm,
    const int *rep, const long long *toff, int *zdst, int *cE, int *csrc)
{
    const int f = tile_filter[blockIdx.x];
    if (!prm[f].active)
        return;
    const BatchFilter F = bf[f];
    const long long k = static_cast<long long>(blockIdx.x) - F.tile0;
    scatter_tile<false>(rep + F.off, F.n, k, toff + 3 * (F.tile0 + f),
        zdst + F.off, cE + F.off, csrc + F.off, nullptr);
}

__global__ void __launch_bounds__(kBlock) k_batch_expand(
    const BatchFilter *bf, const int *tile_filter, const ResampleDev *prm,
    const long long *toff, const int *zdst, const int *cE, const int *csrc,
    float *state, long long ld)
{
    const int f = tile_filter[blockIdx.x];
    if (!prm[f].active)
        return;
    const BatchFilter F = bf[f];
    const long long chunk = static_cast<long long>(blockIdx.x) - F.tile0;
    const long long *to = toff + 3 * (F.tile0 + f);
    const long long K = to[3 * F.ntiles + 0];
    const long long C = to[3 * F.ntiles + 2];
    if (C <= 0 || chunk * kTile >= K)
        return;
    expand_chunk<false, float>(chunk, K, C, zdst + F.off, cE + F.off,
        csrc + F.off, nullptr, state + F.off, 4, ld);
}

int batch_resample(int scheme, int nfilters, const BatchFilter *bf,
    const int *tile_filter, long long ntiles, const ResampleDev *prm,
    const double *l, float *state, long long ld, int *rep, long long *tagg,
    long long *toff, int *zdst, int *cE, int *csrc, const double *tpre,
    TileSeg *segs, double *xstart, int *hard, hipStream_t st)
{
    if (scheme != kSystematic && scheme != kStratified) {
        set_error("batched filters resample with the systematic or stratified scheme");
        return kErrUnsupported;
    }
    const unsigned int gt = static_cast<unsigned int>(ntiles);
    hipLaunchKernelGGL(k_batch_segments, dim3(gt), dim3(kBlock), 0, st, bf,
        tile_filter, prm, l, tpre, segs);
    hipLaunchKernelGGL(k_batch_walk, dim3(nfilters), dim3(kBlock), 0, st, bf,
        prm, l, static_cast<const TileSeg *>(segs), xstart);
    // hard[0] was cleared by this step's k_batch_finalize; one workgroup
    // per listed tile (the list can hold many filters' tiles at once)
    const unsigned int gh = gt;
    if (scheme == kSystematic) {
        hipLaunchKernelGGL(k_batch_count<USys>, dim3(gt), dim3(kBlock), 0, st,
            bf, tile_filter, prm, l, static_cast<const double *>(xstart), rep,
            tagg, hard);
        hipLaunchKernelGGL(k_batch_count_hard<USys>, dim3(gh), dim3(kBlock), 0,
            st, bf, tile_filter, prm, l, static_cast<const double *>(xstart), rep,
            tagg, static_cast<const int *>(hard));
    } else {
        hipLaunchKernelGGL(k_batch_count<UStrat>, dim3(gt), dim3(kBlock), 0, st,
            bf, tile_filter, prm, l, static_cast<const double *>(xstart), rep,
            tagg, hard);
        hipLaunchKernelGGL(k_batch_count_hard<UStrat>, dim3(gh), dim3(kBlock),
            0, st, bf, tile_filter, prm, l, static_cast<const double *>(xstart),
            rep, tagg, static_cast<const int *>(hard));
    }
    hipLaunchKernelGGL(k_batch_scan, dim3(nfilters), dim3(kBlock), 0, st, bf,
        prm, static_cast<const long long *>(tagg), toff);
    hipLaunchKernelGGL(k_batch_scatter, dim3(gt), dim3(kBlock), 0, st, bf,
        tile_filter, prm, static_cast<const int *>(rep),
        static_cast<const long long *>(toff), zdst, cE, csrc);
    hipLaunchKernelGGL(k_batch_expand, dim3(gt), dim3(kBlock), 0, st, bf,
        tile_filter, prm, static_cast<const long long *>(toff),
        static_cast<const int *>(zdst), static_cast<const int *>(cE),
        static_cast<const int *>(csrc), state, ld);
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

// ------------------------------------------------------- event setup

struct HostParams {
    long long M, N;
    unsigned int key;
    unsigned long long word0;
    int scheme;
    int set_weights;   // 1: overwrite active / counts (host-driven)
};

// Host-driven event: fill the device parameters (no host->device copy).
__global__ void k_setup(ResampleDev *p, HostParams h)
{
    p->M = h.M;
    p->nsample = h.N;
    p->ndraw = h.N;
    p->key = h.key;
    p->word0 = h.word0;
    p->active = 1;
    p->err = 0;
    p->words_used = 0;
    p->integ_sum = 0;
    p->equal = 0;
}

// Draw parameters once ndraw is known: delta = fl(1 / ndraw) and the
// systematic start u0 = U * delta drawn first (u01_sequence.hpp:182-187,
// :231-236).
__global__ void k_draw_params(ResampleDev *p, int scheme)
{
    if (!p->active)
        return;
    const long long nd = p->ndraw;
    p->delta = 1 / static_cast<double>(nd);
    if (scheme == kSystematic || scheme == kResidualSystematic) {
        const double U = u01_u32d(philox_word(p->key, p->word0));
        p->u0 = U * p->delta;
    } else {
        p->u0 = 0.0;
    }
}

// Residual: exact total of residuals -> 1/total; R = N - sum(integ).
__global__ void k_resid_params(ResampleDev *p, const double *xstart_total)
{
    if (!p->active)
        return;
    p->resid_total = *xstart_total;
    p->inv_resid = 1 / p->resid_total;
    p->ndraw = p->nsample - p->integ_sum;
}

template <class WSrc>
__global__ void __launch_bounds__(kBlock) k_integ_sum(WSrc src, long long M,
    ResampleDev *p)
{
    if (!p->active)
        return;
    __shared__ long long lds[kWaves];
    long long acc = 0;
    const long long base = static_cast<long long>(blockIdx.x) * kTile;
    // clamped, unconditional loads, all in flight before their first use
    long long v[kItems];
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const long long g = base + j * kBlock + threadIdx.x;
        v[j] = src.integ(g < M ? g : M - 1);
    }
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const long long g = base + j * kBlock + threadIdx.x;
        keep_loaded(v[j]);
        acc += g < M ? v[j] : 0;
    }
    acc = block_reduce(acc, OpAddL(), lds);
    if (threadIdx.x == 0)
        atomicAdd(reinterpret_cast<unsigned long long *>(&p->integ_sum),
            static_cast<unsigned long long>(acc));
}

// Words of the resampling stream one event consumes: the lazy sequences
// draw one word per new index up to the first failing comparison
// (transform.hpp:61-69), systematic one word in its constructor.
__global__ void k_words_used(ResampleDev *p, const int *rep, int scheme,
    int resid, const double *W, int wkind, const double *lw)
{
    if (!p->active)
        return;
    const long long M = p->M;
    const long long nd = p->ndraw;
    unsigned long long used = 0;
    if (scheme == kSystematic || scheme == kResidualSystematic) {
        used = 1;
    } else if (M >= 2 && nd >= 1) {
        long long last = rep[M - 1];
        if (resid) {
            double w = wkind == kWeightLog ?
                (p->equal ? 1.0 / static_cast<double>(M) :
                            p->inv_sum * wexp(lw[M - 1] - p->lmax)) :
                W[M - 1];
            double ip;
            modf(static_cast<double>(p->nsample) * w, &ip);
            last -= static_cast<long long>(ip);
        }
        const long long jend = nd - last;
        used = static_cast<unsigned long long>(jend + 1 < nd ? jend + 1 : nd);
    }
    p->words_used = used;
}

#include "fast_event.hpp"

// ------------------------------------------------------------ host side

// The last error of the calling host thread.
static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

const char *get_error() { return g_err.c_str(); }

bool Workspace::ensure(long long n)
{
    const long long tiles = workspace_tiles(n);
    if (n <= cap_n && tiles <= cap_tiles && prm)
        return true;
    if (fixed) {
        set_error("size " + std::to_string(n) +
            " exceeds the fixed workspace (at most " + std::to_string(cap_n) +
            " weights)");
        return false;
    }
    release();
    const long long nn = n < 1 ? 1 : n;
    const long long ngroups = tiles / kGroup + 2;
    bool ok = true;
    auto A = [&](auto **p, size_t bytes) {
        void *q = nullptr;
        if (hipMalloc(&q, bytes ? bytes : 8) != hipSuccess || q == nullptr)
            ok = false;
        *p = static_cast<std::remove_reference_t<decltype(**p)> *>(q);
    };
    A(&tsum, sizeof(double) * tiles);
    A(&tpre, sizeof(double) * tiles);
    A(&segs, sizeof(TileSeg) * tiles);
    A(&xstart, sizeof(double) * (tiles + 1));
    A(&xgstart, sizeof(double) * (ngroups + 1));
    A(&gmode, sizeof(int) * (ngroups + 1));
    A(&groups, sizeof(GroupSeg) * (ngroups + 1));
    A(&tagg, sizeof(long long) * 3 * tiles);
    A(&toff, sizeof(long long) * 3 * (tiles + 1));
    A(&zdst, sizeof(int) * nn);
    A(&cE, sizeof(int) * nn);
    A(&csrc, sizeof(int) * nn);
    A(&tq, sizeof(int) * 2 * tiles);
    A(&v, sizeof(double) * nn);
    A(&aux, sizeof(double) * nn);
    A(&rep, sizeof(int) * nn);
    A(&hard, sizeof(int) * (tiles + 1));
    A(&fsum, sizeof(FeSum) * tiles);
    A(&fcross, sizeof(FeCross) * tiles);
    A(&fpred, sizeof(int) * (tiles + 1));
    A(&prm, sizeof(ResampleDev));
    A(&scal, sizeof(long long) * 16);
    if (!ok || !all_allocated()) {
        release();
        set_error("workspace allocation failed");
        return false;
    }
    // no predicted tiles until a finalize lists them (a count above the
    // list's capacity: every tile in order)
    if (hipMemset(fpred, 0x7F, sizeof(int)) != hipSuccess) {
        release();
        set_error("workspace initialisation failed");
        return false;
    }
    cap_n = nn;
    cap_tiles = tiles;
    return true;
}

bool Workspace::all_allocated() const
{
    const void *ps[] = {tsum, tpre, segs, xstart, xgstart, gmode, groups, tagg,
        toff, zdst, cE, csrc, tq, v, aux, rep, hard, fsum, fcross, fpred, prm, scal};
    for (const void *p : ps)
        if (!p)
            return false;
    return true;
}

void Workspace::release()
{
    if (fixed)
        return;
    void *ps[] = {tsum, tpre, segs, xstart, xgstart, gmode, groups, tagg, toff,
        zdst, cE, csrc, tq, v, aux, rep, hard, fsum, fcross, fpred, prm, scal};
    for (void *p : ps)
        if (p)
            (void)hipFree(p);
    tsum = tpre = xstart = xgstart = v = aux = nullptr;
    segs = nullptr;
    gmode = zdst = cE = csrc = tq = rep = hard = nullptr;
    fsum = nullptr;
    fcross = nullptr;
    fpred = nullptr;
    groups = nullptr;
    tagg = toff = scal = nullptr;
    prm = nullptr;
    cap_n = cap_tiles = 0;
}

namespace {

// Workspaces of exited host threads, per device. Never destroyed: a
// thread-exit handler may run during process exit.
struct WsPool {
    std::mutex mu;
    std::vector<std::vector<Workspace *>> spare;
};

WsPool &ws_pool()
{
    static WsPool *p = new WsPool();
    return *p;
}

// A host thread's workspaces; at thread exit they go back to the pool (no
// device call here: the runtime may already be shutting down).
struct ThreadWs {
    std::vector<Workspace *> ws;
    ~ThreadWs()
    {
        WsPool &pool = ws_pool();
        std::lock_guard<std::mutex> lk(pool.mu);
        for (size_t d = 0; d < ws.size(); ++d) {
            if (!ws[d])
                continue;
            if (pool.spare.size() <= d)
                pool.spare.resize(d + 1);
            pool.spare[d].push_back(ws[d]);
        }
    }
};

} // namespace

Workspace &default_workspace(int device)
{
    thread_local ThreadWs mine;
    if (device < 0)
        device = 0;
    const size_t d = static_cast<size_t>(device);
    if (mine.ws.size() <= d)
        mine.ws.resize(d + 1, nullptr);
    if (!mine.ws[d]) {
        Workspace *w = nullptr;
        {
            WsPool &pool = ws_pool();
            std::lock_guard<std::mutex> lk(pool.mu);
            if (pool.spare.size() > d && !pool.spare[d].empty()) {
                w = pool.spare[d].back();
                pool.spare[d].pop_back();
            }
        }
        if (w)
            (void)hipDeviceSynchronize();   // the old owner's queued work
        else
            w = new Workspace();
        mine.ws[d] = w;
    }
    return *mine.ws[d];
}

static inline unsigned int grid_tiles(long long n)
{
    const long long t = (n + kTile - 1) / kTile;
    return static_cast<unsigned int>(t < 1 ? 1 : t);
}

// Runs the fold passes (tile sums, approximate scan, segments, groups,
// resolver) over n weights from src; n may live on the device (ndev).
template <class WSrc>
static int run_fold(Workspace &ws, WSrc src, long long nmax,
    const long long *ndev, bool tpre_ready = false)
{
    hipStream_t st = ws.stream;
    const unsigned int gt = grid_tiles(nmax);
    const unsigned int gg = static_cast<unsigned int>((gt + kGroup - 1) / kGroup);
    int *err = reinterpret_cast<int *>(&ws.prm->err);
    // tpre_ready: the approximate tile prefix was written from the weight
    // pass's block partials (blocks and tiles both hold kTile particles)
    if (!tpre_ready) {
        hipLaunchKernelGGL(k_tile_sum<WSrc>, dim3(gt), dim3(kBlock), 0, st,
            src, nmax, ndev, ws.prm, ws.tsum, err);
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kBlock), 0, st,
            ws.tsum, nmax, ndev, ws.prm, ws.tpre);
    }
    hipLaunchKernelGGL(k_segments<WSrc>, dim3(gt), dim3(kBlock), 0, st, src,
        nmax, ndev, ws.prm, ws.tpre, ws.segs);
    hipLaunchKernelGGL(k_groups, dim3(gg), dim3(kWave), 0, st, nmax, ndev,
        ws.prm, ws.segs, ws.groups);
    hipLaunchKernelGGL(k_resolve<WSrc>, dim3(1), dim3(kBlock), 0, st, src,
        nmax, ndev, ws.prm, ws.segs, ws.groups, ws.xstart, ws.xgstart,
        ws.gmode);
    hipLaunchKernelGGL(k_starts, dim3(gg), dim3(kWave), 0, st, nmax, ndev,
        ws.prm, ws.segs, ws.xgstart, ws.gmode, ws.xstart);
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

static FoldBufs fold_bufs(Workspace &ws)
{
    return FoldBufs{ws.xstart};
}

template <class WSrc, class USrc>
static int run_count(Workspace &ws, WSrc src, USrc us, long long M, int *rep,
    bool tpre_ready = false)
{
    int rc = run_fold(ws, src, M, nullptr, tpre_ready);
    if (rc != kOk)
        return rc;
    // one block per tile; a tile whose sums cross a binade is folded in
    // full by its own block while the others take the one-binade fast path
    const unsigned int gt = grid_tiles(M);
    hipLaunchKernelGGL((k_count_all<WSrc, USrc>), dim3(gt), dim3(kBlock), 0,
        ws.stream, src, us, M, static_cast<const long long *>(nullptr), ws.prm,
        fold_bufs(ws), rep, ws.tagg);
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

// The counts of a particle set's event (fast_event.hpp): log-weights,
// systematic or stratified draws prepared on the device (ws.tpre and
// ws.fpred written by the finalize), M == N.
template <class WSrc, class USrc>
static int fast_count(Workspace &ws, WSrc wb, USrc us, long long M, int *rep)
{
    hipStream_t st = ws.stream;
    const unsigned int gt = grid_tiles(M);
    const long long *none = nullptr;
    hipLaunchKernelGGL(k_fe_summ<WSrc>, dim3(kFePred + gt), dim3(kBlock), 0, st, wb,
        M, none, ws.prm, ws.tpre, ws.fpred, ws.segs, ws.fsum, ws.fcross);
    hipLaunchKernelGGL(k_fe_resolve<WSrc>, dim3(1), dim3(kFeThreads), 0, st, wb, M,
        none, ws.prm, ws.segs, ws.fsum, ws.fcross, ws.groups, ws.xstart,
        ws.xgstart, ws.gmode);
    if constexpr (std::is_same<USrc, UArr>::value)
        hipLaunchKernelGGL(k_fe_arr_bounds, dim3((gt + kWaves - 1) / kWaves),
            dim3(kBlock), 0, st, us, M, ws.prm, ws.xstart, ws.hard);
    if constexpr (std::is_same<USrc, UStrat>::value)
        hipLaunchKernelGGL(k_fe_strat_bounds, dim3((gt + kBlock - 1) / kBlock),
            dim3(kBlock), 0, st, us, M, ws.prm, ws.xstart, ws.hard);
    hipLaunchKernelGGL((k_fe_count<WSrc, USrc>), dim3(gt), dim3(kBlock), 0, st,
        wb, us, M, ws.prm, ws.xstart, rep, ws.tagg, ws.csrc, ws.cE,
        static_cast<const int *>(ws.hard));
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

static int check_err(Workspace &ws)
{
    int err = 0;
    VSMCH_HIP(hipMemcpyAsync(&err, &ws.prm->err, sizeof(int),
        hipMemcpyDeviceToHost, ws.stream));
    VSMCH_HIP(hipStreamSynchronize(ws.stream));
    if (err) {
        set_error("weights must be non-negative and not NaN");
        return kErrWeights;
    }
    return kOk;
}

static int setup_host(Workspace &ws, long long M, long long N,
    unsigned int key, unsigned long long word0, int scheme)
{
    HostParams h{M, N, key, word0, scheme, 1};
    hipLaunchKernelGGL(k_setup, dim3(1), dim3(1), 0, ws.stream, ws.prm, h);
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

// Running maximum of u (n entries) into ws.v.
static int run_maxscan(Workspace &ws, const double *u, long long n,
    const long long *ndev)
{
    const unsigned int gt = grid_tiles(n);
    hipLaunchKernelGGL(k_tile_max, dim3(gt), dim3(kBlock), 0, ws.stream, u, n,
        ndev, ws.prm, ws.tsum);
    hipLaunchKernelGGL(k_scan_max, dim3(1), dim3(kBlock), 0, ws.stream,
        ws.tsum, n, ndev, ws.prm, ws.tpre);
    hipLaunchKernelGGL(k_apply_max, dim3(gt), dim3(kBlock), 0, ws.stream, u,
        n, ndev, ws.prm, ws.tpre, ws.v);
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

// The sorted draws of a particle set's multinomial event (n draws of the
// stream prm->key, prm->word0; *ndev when given, at most n) into out with
// the fast event's passes
// (fast_event.hpp); the terms in ws.v, their tile prefix in ws.tsum and
// predicted crossings in ws.tq. Leaves ws.tpre and ws.fpred (the weights')
// untouched; the running maximum runs in place only where u decreased.
static int run_sorted_fast(Workspace &ws, long long n, const long long *ndev,
    double *out)
{
    hipStream_t st = ws.stream;
    const unsigned int gt = grid_tiles(n);
    hipLaunchKernelGGL(k_fe_sorted_terms, dim3(gt), dim3(kBlock), 0, st, n, ndev,
        ws.prm, ws.v, ws.tsum);
    hipLaunchKernelGGL(k_fe_sorted_scan, dim3(1), dim3(kFeScanThreads), 0, st, n,
        ndev, ws.prm, ws.tsum, ws.tq);
    hipLaunchKernelGGL(k_fe_summ<WArray>, dim3(kFePred + gt), dim3(kBlock), 0, st,
        WArray{ws.v}, n, ndev, ws.prm, ws.tsum, ws.tq, ws.segs, ws.fsum, ws.fcross);
    hipLaunchKernelGGL(k_fe_resolve<WArray>, dim3(1), dim3(kFeThreads), 0, st,
        WArray{ws.v}, n, ndev, ws.prm, ws.segs, ws.fsum, ws.fcross, ws.groups,
        ws.xstart, ws.xgstart, ws.gmode);
    hipLaunchKernelGGL(k_fe_sorted_prefix, dim3(gt), dim3(kBlock), 0, st, n, ndev,
        ws.prm, ws.v, ws.xstart, out, ws.tsum);
    hipLaunchKernelGGL(k_fe_sorted_max, dim3(gt), dim3(kBlock), 0, st, n, ndev,
        ws.prm, ws.tsum, out);
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

int trans_u01_rep(Workspace &ws, long long M, long long N, const double *W,
    const double *u01, int *rep)
{
    if (M <= 0)
        return kOk;
    if (!ws.ensure(M > N ? M : N))
        return kErrAlloc;
    int rc = setup_host(ws, M, N, 0, 0, kMultinomial);
    if (rc != kOk)
        return rc;
    if (M == 1) {
        hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, ws.stream, rep,
            1LL, static_cast<int>(N), ws.prm);
        VSMCH_HIP(hipGetLastError());
        return kOk;
    }
    if (N == 0) {
        hipLaunchKernelGGL(k_fill<int>, dim3((M + 255) / 256), dim3(256), 0,
            ws.stream, rep, M, 0, ws.prm);
        VSMCH_HIP(hipGetLastError());
        return kOk;
    }
    rc = run_maxscan(ws, u01, N, nullptr);
    if (rc != kOk)
        return rc;
    rc = run_count(ws, WArray{W}, UArr{ws.v, ws.prm}, M, rep);
    if (rc != kOk)
        return rc;
    return check_err(ws);
}

int exact_cumsum(Workspace &ws, long long n, const double *w, double *out)
{
    if (n <= 0)
        return kOk;
    if (!ws.ensure(n))
        return kErrAlloc;
    int rc = setup_host(ws, n, n, 0, 0, kMultinomial);
    if (rc != kOk)
        return rc;
    rc = run_fold(ws, WArray{w}, n, nullptr);
    if (rc != kOk)
        return rc;
    hipLaunchKernelGGL((k_prefix<WArray, EpiRaw>), dim3(grid_tiles(n)),
        dim3(kBlock), 0, ws.stream, WArray{w}, n,
        static_cast<const long long *>(nullptr), ws.prm, fold_bufs(ws),
        EpiRaw{out});
    VSMCH_HIP(hipGetLastError());
    return check_err(ws);
}

// Sorted uniforms of U01SequenceSorted for ndraw draws into out; ndraw may
// live on the device (ndev, with nmax an upper bound).
static int run_sorted(Workspace &ws, long long nmax, const long long *ndev,
    double *out)
{
    WSortedTerms terms{ws.prm};
    int rc = run_fold(ws, terms, nmax, ndev);
    if (rc != kOk)
        return rc;
    hipLaunchKernelGGL((k_prefix<WSortedTerms, EpiSorted>),
        dim3(grid_tiles(nmax)), dim3(kBlock), 0, ws.stream, terms, nmax, ndev,
        ws.prm, fold_bufs(ws), EpiSorted{out});
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

int sorted_uniforms(Workspace &ws, long long ndraw, unsigned int key,
    unsigned long long word0, double *out)
{
    if (ndraw <= 0)
        return kOk;
    if (!ws.ensure(ndraw))
        return kErrAlloc;
    int rc = setup_host(ws, ndraw, ndraw, key, word0, kMultinomial);
    if (rc != kOk)
        return rc;
    rc = run_sorted(ws, ndraw, nullptr, out);
    if (rc != kOk)
        return rc;
    return check_err(ws);
}

// u01_sorted over the caller's uniforms (u01_sequence.hpp:54-62): the
// terms -log(u_i) / (N - i) folded in order, s_i -> 1 - exp(-s_i) (the
// reference's lmax = -s_i, rounded the same way by symmetry).
struct WLogUArr {
    const double *u;
    long long n;
    __device__ double operator()(long long k) const
    {
        return -(log(u[k]) / static_cast<double>(n - k));
    }
    __device__ long long integ(long long) const { return 0; }
    __device__ WLogUArr loaded() const { return *this; }
};

// u01_systematic (u01_sequence.hpp:80-89): u_0 = u delta and u_i = u_{i-1}
// + delta, the sequential sums of (u delta, delta, delta, ...).
struct WSysSteps {
    double first;
    double delta;
    __device__ double operator()(long long k) const { return k == 0 ? first : delta; }
    __device__ long long integ(long long) const { return 0; }
    __device__ WSysSteps loaded() const { return *this; }
};

// u01_stratified (u01_sequence.hpp:67-74): u_i delta + i delta, no FMA;
// delta = 1 / n rounded on the host.
__global__ void k_u01_stratified(long long n, double delta, const double *u,
    double *out)
{
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    out[i] = rn_add(rn_mul(u[i], delta), rn_mul(static_cast<double>(i), delta));
}

template <class WSrc, class Epi>
static int fold_prefix(Workspace &ws, long long n, WSrc src, Epi epi)
{
    if (!ws.ensure(n))
        return kErrAlloc;
    int rc = setup_host(ws, n, n, 0, 0, kMultinomial);
    if (rc != kOk)
        return rc;
    rc = run_fold(ws, src, n, nullptr);
    if (rc != kOk)
        return rc;
    hipLaunchKernelGGL((k_prefix<WSrc, Epi>), dim3(grid_tiles(n)), dim3(kBlock),
        0, ws.stream, src, n, static_cast<const long long *>(nullptr), ws.prm,
        fold_bufs(ws), epi);
    VSMCH_HIP(hipGetLastError());
    return check_err(ws);
}

int u01_sorted_seq(Workspace &ws, long long n, const double *u, double *out)
{
    if (n <= 0)
        return kOk;
    return fold_prefix(ws, n, WLogUArr{u, n}, EpiSorted{out});
}

int u01_stratified_seq(Workspace &ws, long long n, const double *u, double *out)
{
    if (n <= 0)
        return kOk;
    const double delta = 1 / static_cast<double>(n);
    hipLaunchKernelGGL(k_u01_stratified, dim3(static_cast<unsigned>((n + 255) / 256)),
        dim3(256), 0, ws.stream, n, delta, u, out);
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

int u01_systematic_seq(Workspace &ws, long long n, double u, double *out)
{
    if (n <= 0)
        return kOk;
    const double delta = 1.0 / static_cast<double>(n);
    return fold_prefix(ws, n, WSysSteps{u * delta, delta}, EpiRaw{out});
}

template <class Base>
static int scheme_counts(Workspace &ws, int scheme, long long M, long long N,
    Base wb, int *rep, bool prepared)
{
    hipStream_t st = ws.stream;
    const long long *ndraw_dev = &ws.prm->ndraw;
    const bool resid = scheme == kResidual || scheme == kResidualStratified ||
        scheme == kResidualSystematic;
    int rc;
    if (prepared && !resid) {
        // draw parameters and the weights' tile prefix were set on the device
        if (scheme == kMultinomial) {
            if constexpr (std::is_same<Base, WLog>::value) {
                if (M == N) {
                    rc = run_sorted_fast(ws, N, nullptr, ws.aux);
                    if (rc != kOk)
                        return rc;
                    return fast_count(ws, wb, UArr{ws.aux, ws.prm}, M, rep);
                }
            }
            // the sorted draws first, their fold's tile prefix in the summary
            // buffer (free until the weights' summary pass; ws.tpre holds the
            // weights' prefix for that pass)
            Workspace wd = ws;
            wd.tpre = reinterpret_cast<double *>(ws.fsum);
            rc = run_sorted(wd, N, ndraw_dev, ws.aux);
            if (rc != kOk)
                return rc;
            rc = run_maxscan(wd, ws.aux, N, ndraw_dev);
            if (rc != kOk)
                return rc;
            return run_count(ws, wb, UArr{ws.v, ws.prm}, M, rep, true);
        }
        if constexpr (std::is_same<Base, WLog>::value) {
            if (M == N) {
                if (scheme == kStratified)
                    return fast_count(ws, wb, UStrat{ws.prm}, M, rep);
                return fast_count(ws, wb, USys{ws.prm}, M, rep);
            }
        }
        if (scheme == kStratified)
            return run_count(ws, wb, UStrat{ws.prm}, M, rep, true);
        return run_count(ws, wb, USys{ws.prm}, M, rep, true);
    }
    if (!resid) {
        hipLaunchKernelGGL(k_draw_params, dim3(1), dim3(1), 0, st, ws.prm,
            scheme);
        if (scheme == kMultinomial) {
            rc = run_sorted(ws, N, ndraw_dev, ws.aux);
            if (rc != kOk)
                return rc;
            rc = run_maxscan(ws, ws.aux, N, ndraw_dev);
            if (rc != kOk)
                return rc;
            return run_count(ws, wb, UArr{ws.v, ws.prm}, M, rep);
        }
        if (scheme == kStratified)
            return run_count(ws, wb, UStrat{ws.prm}, M, rep);
        return run_count(ws, wb, USys{ws.prm}, M, rep);
    }
    // residual schemes: exact total of the raw residuals, integer parts
    const double coeff = static_cast<double>(N);
    if constexpr (std::is_same<Base, WLog>::value) {
        if (prepared && M == N) {
            // a particle set's event: the raw residuals' tile sums with the
            // integer parts', their exact total from the fast event's
            // summaries and resolver and the last tile's fold, the residual
            // parameters and the normalised residuals' prefix, then the fast
            // event's counting passes (R sorted draws first for the
            // multinomial residual scheme)
            const unsigned int gt = grid_tiles(M);
            const long long *none = nullptr;
            hipLaunchKernelGGL(k_fe_resid_tiles, dim3(gt), dim3(kBlock), 0, st, wb,
                M, coeff, ws.prm, ws.tsum, ws.tagg);
            // raw residuals: prefix in ws.tsum, predicted crossings in ws.tq
            hipLaunchKernelGGL(k_fe_sorted_scan, dim3(1), dim3(kFeScanThreads), 0,
                st, M, none, ws.prm, ws.tsum, ws.tq);
            WResidRaw<Base> raw{wb, coeff};
            hipLaunchKernelGGL(k_fe_summ<WResidRaw<Base>>, dim3(kFePred + gt),
                dim3(kBlock), 0, st, raw, M, none, ws.prm, ws.tsum, ws.tq,
                ws.segs, ws.fsum, ws.fcross);
            hipLaunchKernelGGL(k_fe_resolve<WResidRaw<Base>>, dim3(1),
                dim3(kFeThreads), 0, st, raw, M, none, ws.prm, ws.segs, ws.fsum,
                ws.fcross, ws.groups, ws.xstart, ws.xgstart, ws.gmode);
            hipLaunchKernelGGL(k_fe_resid_params<Base>, dim3(1), dim3(kBlock), 0,
                st, raw, M, ws.prm, ws.xstart, ws.tagg, ws.tsum, ws.tpre,
                ws.fpred, scheme);
            VSMCH_HIP(hipGetLastError());
            WResid<Base> res{wb, coeff, ws.prm};
            if (scheme == kResidual) {
                rc = run_sorted_fast(ws, N, ndraw_dev, ws.aux);
                if (rc != kOk)
                    return rc;
                return fast_count(ws, res, UArr{ws.aux, ws.prm}, M, rep);
            }
            if (scheme == kResidualStratified)
                return fast_count(ws, res, UStrat{ws.prm}, M, rep);
            return fast_count(ws, res, USys{ws.prm}, M, rep);
        }
    }
    WResidRaw<Base> raw{wb, coeff};
    rc = run_fold(ws, raw, M, nullptr);
    if (rc != kOk)
        return rc;
    hipLaunchKernelGGL(k_integ_sum<WResidRaw<Base>>, dim3(grid_tiles(M)),
        dim3(kBlock), 0, st, raw, M, ws.prm);
    const long long ntiles = (M + kTile - 1) / kTile;
    hipLaunchKernelGGL(k_resid_params, dim3(1), dim3(1), 0, st, ws.prm,
        ws.xstart + ntiles);
    hipLaunchKernelGGL(k_draw_params, dim3(1), dim3(1), 0, st, ws.prm, scheme);
    VSMCH_HIP(hipGetLastError());
    WResid<Base> res{wb, coeff, ws.prm};
    if (scheme == kResidual) {
        rc = run_sorted(ws, N, ndraw_dev, ws.aux);
        if (rc != kOk)
            return rc;
        rc = run_maxscan(ws, ws.aux, N, ndraw_dev);
        if (rc != kOk)
            return rc;
        return run_count(ws, res, UArr{ws.v, ws.prm}, M, rep);
    }
    if (scheme == kResidualStratified)
        return run_count(ws, res, UStrat{ws.prm}, M, rep);
    return run_count(ws, res, USys{ws.prm}, M, rep);
}

int resample_scheme(Workspace &ws, int scheme, long long M, long long N,
    unsigned int key, unsigned long long word0, const double *weights,
    int weight_kind, int *rep, bool device_active, bool prepared)
{
    if (M <= 0)
        return kOk;
    if (!ws.ensure(M > N ? M : N))
        return kErrAlloc;
    int rc;
    if (!device_active) {
        rc = setup_host(ws, M, N, key, word0, scheme);
        if (rc != kOk)
            return rc;
    }
    if (M == 1 || N == 0) {
        // transform.hpp:49-59: nothing is drawn except the systematic
        // sequence's constructor word.
        if (M == 1)
            hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, ws.stream,
                rep, 1LL, static_cast<int>(N), ws.prm);
        else
            hipLaunchKernelGGL(k_fill<int>, dim3((M + 255) / 256), dim3(256),
                0, ws.stream, rep, M, 0, ws.prm);
        const unsigned long long used =
            (scheme == kSystematic || scheme == kResidualSystematic) ? 1 : 0;
        hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(1), dim3(1), 0,
            ws.stream, &ws.prm->words_used, 1LL, used, ws.prm);
        VSMCH_HIP(hipGetLastError());
        return kOk;
    }
    const bool prep = prepared && device_active && fast_event_scheme(scheme);
    if (weight_kind == kWeightLog)
        rc = scheme_counts(ws, scheme, M, N, WLog{weights, ws.prm}, rep, prep);
    else
        rc = scheme_counts(ws, scheme, M, N, WArray{weights}, rep, prep);
    if (rc != kOk)
        return rc;
    const bool resid = scheme == kResidual || scheme == kResidualStratified ||
        scheme == kResidualSystematic;
    // the fast event's counting pass counts the stream words itself
    if (!prep)
        hipLaunchKernelGGL(k_words_used, dim3(1), dim3(1), 0, ws.stream, ws.prm,
            rep, scheme, resid ? 1 : 0,
            weight_kind == kWeightArray ? weights : nullptr, weight_kind,
            weight_kind == kWeightLog ? weights : nullptr);
    VSMCH_HIP(hipGetLastError());
    if (!device_active)
        return check_err(ws);
    return kOk;
}

int read_words_used(Workspace &ws, unsigned long long *out)
{
    VSMCH_HIP(hipMemcpyAsync(out, &ws.prm->words_used, sizeof(*out),
        hipMemcpyDeviceToHost, ws.stream));
    VSMCH_HIP(hipStreamSynchronize(ws.stream));
    return kOk;
}

int trans_rep_index(Workspace &ws, long long M, long long N, const int *rep,
    int *idx, void *state, int dim, long long ld, int elem_bytes,
    bool device_active, bool force_expand, bool agg_ready, const EventEnd &end,
    bool tile_sources)
{
    if (M <= 0 || N <= 0)
        return kOk;
    if (!ws.ensure(M > N ? M : N))
        return kErrAlloc;
    hipStream_t st = ws.stream;
    ResampleDev *prm = device_active ? ws.prm : nullptr;
    const long long ntiles = (M + kTile - 1) / kTile;
    const bool expand = M != N || force_expand;
    if (expand) {
        hipLaunchKernelGGL(k_ridx_agg<true>, dim3(ntiles), dim3(kBlock), 0,
            st, rep, M, prm, ws.tagg);
    } else if (!agg_ready) {
        hipLaunchKernelGGL(k_ridx_agg<false>, dim3(ntiles), dim3(kBlock), 0,
            st, rep, M, prm, ws.tagg);
    }
    if (expand) {
        hipLaunchKernelGGL(k_ridx_scan, dim3(1), dim3(kScanThreads), 0, st,
            ws.tagg, ntiles, prm, ws.toff);
        hipLaunchKernelGGL(k_ridx_scatter<true>, dim3(ntiles), dim3(kBlock), 0,
            st, rep, M, prm, ws.toff, ws.zdst, ws.cE, ws.csrc, idx);
        // the gather's block 0 ends the event (one launch fewer per step)
        hipLaunchKernelGGL((k_expand<true, float>), dim3(grid_tiles(N)),
            dim3(kBlock), 0, st, prm, ws.toff, ntiles, ws.zdst, ws.cE, ws.csrc,
            idx, static_cast<float *>(nullptr), 0, 0LL, ws.prm, end);
        VSMCH_HIP(hipGetLastError());
        return kOk;
    }
    // M == N: the prefixes (and each tile's slice of sources), the
    // compacted sources, then every tile's indices (and gathered rows) in
    // one pass; its block 0 ends the event
    // the tiles find their slice of sources themselves (two wave searches
    // over the prefixes, in flight with their rep loads)
    hipLaunchKernelGGL(k_ridx_prefix, dim3(1), dim3(kScanThreads), 0, st,
        ws.tagg, ntiles, prm, ws.toff, static_cast<int *>(nullptr));
    const int *tq = nullptr;
    if (tile_sources && agg_ready) {
        if (state == nullptr || elem_bytes == 4) {
            hipLaunchKernelGGL((k_ridx_fill<float, true>), dim3(ntiles), dim3(kBlock), 0,
                st, prm, rep, M, ws.toff, tq, ntiles, ws.cE, ws.csrc, idx,
                static_cast<float *>(state), dim, ld, ws.prm, end);
        } else {
            hipLaunchKernelGGL((k_ridx_fill<double, true>), dim3(ntiles), dim3(kBlock),
                0, st, prm, rep, M, ws.toff, tq, ntiles, ws.cE, ws.csrc, idx,
                static_cast<double *>(state), dim, ld, ws.prm, end);
        }
        VSMCH_HIP(hipGetLastError());
        return kOk;
    }
    hipLaunchKernelGGL(k_ridx_sources, dim3(ntiles), dim3(kBlock), 0, st, rep,
        M, prm, ws.toff, ws.cE, ws.csrc);
    if (state == nullptr || elem_bytes == 4) {
        hipLaunchKernelGGL((k_ridx_fill<float>), dim3(ntiles), dim3(kBlock), 0,
            st, prm, rep, M, ws.toff, tq, ntiles, ws.cE, ws.csrc, idx,
            static_cast<float *>(state), dim, ld, ws.prm, end);
    } else {
        hipLaunchKernelGGL((k_ridx_fill<double>), dim3(ntiles), dim3(kBlock),
            0, st, prm, rep, M, ws.toff, tq, ntiles, ws.cE, ws.csrc, idx,
            static_cast<double *>(state), dim, ld, ws.prm, end);
    }
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

int trans_index_rep(Workspace &ws, long long M, long long N, const int *idx,
    int *rep)
{
    if (M <= 0 || N <= 0)
        return kOk;
    hipStream_t st = ws.stream;
    hipLaunchKernelGGL(k_fill<int>, dim3((M + 255) / 256), dim3(256), 0, st,
        rep, M, 0, static_cast<const ResampleDev *>(nullptr));
    hipLaunchKernelGGL(k_index_rep, dim3((N + 255) / 256), dim3(256), 0, st,
        idx, N, rep);
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

__global__ void k_resid_write(const double *W, long long M, double coeff,
    double *resid, int *integ)
{
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < M) {
        double ip;
        resid[i] = modf(coeff * W[i], &ip);
        integ[i] = static_cast<int>(ip);
    }
}

__global__ void k_resid_scale(double *resid, long long M, const ResampleDev *p)
{
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < M)
        resid[i] = p->inv_resid * resid[i];
}

int trans_residual(Workspace &ws, long long M, long long N, const double *W,
    double *resid, int *integ, long long *R_out)
{
    if (M <= 0) {
        *R_out = N;
        return kOk;
    }
    if (!ws.ensure(M > N ? M : N))
        return kErrAlloc;
    int rc = setup_host(ws, M, N, 0, 0, kResidual);
    if (rc != kOk)
        return rc;
    hipStream_t st = ws.stream;
    const double coeff = static_cast<double>(N);
    hipLaunchKernelGGL(k_resid_write, dim3((M + 255) / 256), dim3(256), 0, st,
        W, M, coeff, resid, integ);
    WResidRaw<WArray> raw{WArray{W}, coeff};
    rc = run_fold(ws, raw, M, nullptr);
    if (rc != kOk)
        return rc;
    hipLaunchKernelGGL(k_integ_sum<WResidRaw<WArray>>, dim3(grid_tiles(M)),
        dim3(kBlock), 0, st, raw, M, ws.prm);
    const long long ntiles = (M + kTile - 1) / kTile;
    hipLaunchKernelGGL(k_resid_params, dim3(1), dim3(1), 0, st, ws.prm,
        ws.xstart + ntiles);
    hipLaunchKernelGGL(k_resid_scale, dim3((M + 255) / 256), dim3(256), 0, st,
        resid, M, ws.prm);
    VSMCH_HIP(hipGetLastError());
    long long isum = 0;
    VSMCH_HIP(hipMemcpyAsync(&isum, &ws.prm->integ_sum, sizeof(isum),
        hipMemcpyDeviceToHost, st));
    rc = check_err(ws);
    if (rc != kOk)
        return rc;
    *R_out = N - isum;
    return kOk;
}

} // namespace vsmch
