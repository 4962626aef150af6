// fast_event.hpp -- replication counts of a particle set's resampling event:
// weights from the log-weights (W_i = e^(l_i - lmax) / sum), systematic or
// stratified uniforms, M == N. Included by resample.hip after the fold
// passes it falls back on, inside namespace vsmch.
//
// The sequential cumulative weights of resample_trans_u01_rep
// (include/vsmc/resample/transform.hpp:61-69) are reproduced exactly, as in
// fold.hpp (within one binade a sum is an integer multiple of the binade's
// ulp and adding w adds the rounded step; a sum that leaves the binade is
// one real FP64 add):
//  * k_fe_summ summarises every tile from its approximate start tpre[t]
//    (the move's block partials): a tile whose approximate sums stay in one
//    binade without half-ulp ties by the integer total of its rounded steps,
//    from registers; a tile with one crossing of the approximate sums and no
//    ties by the totals before and after the crossing and the weight that
//    crosses; any other (tile 0, whose sums start at 0) by the full summary
//    of fold.hpp (runs, their maps and the crossing weights). Those take a
//    block tens of microseconds, so the tiles predicted to need them (from
//    tpre alone, listed by the finalize) run in the grid's first blocks,
//    beside the others.
//  * k_fe_resolve (one workgroup): consecutive one-binade tiles form a
//    segment whose map is the composition of theirs (a segmented block
//    scan); one thread walks the segments and the other tiles in order,
//    verifying every predicted binade against the exact values, and every
//    tile's exact start follows. A failed verification falls back, in the
//    same workgroup, to the general resolver of the fold passes.
//  * k_fe_count: every tile's sums from its exact start: one block scan of
//    the rounded steps when the tile stays in the start's binade without
//    ties; otherwise an FP64 block scan with a proven error bound decides
//    every count unless a draw lies within the bound of a sum, and only then
//    is the tile folded exactly by fold.hpp's passes.
#pragma once

constexpr int kFeSegs = 16;            // full summaries a resolver chunk holds
constexpr int kFePer = 16;             // resolver: tiles per thread per chunk
constexpr long long kFeChunk = static_cast<long long>(kBlock) * kFePer;
constexpr int kFeChainMax = 512;       // chain items per chunk
constexpr int kFeListTag = -100000;    // chain item tags below: full summaries
constexpr int kFeCrossTag = -200000;   // and below these: one-crossing tiles

// Tile summary kinds.
enum {
    kFeMap = 0,     // sums stay in binade e: map (d0, d1) by the start's parity
    kFeList = 1,    // full summary in segs[t] (fold::apply_segments)
    kFeCross = 2,   // binade e (total d0) up to one weight, then FeCross
};

struct FeSum {
    int kind;
    int e;
    long long d0;
    long long d1;
};

// The rest of a one-crossing tile: the crossing weight wc, then binade e2
// with total d2.
struct FeCross {
    int e2;
    int pad;
    long long d2;
    double wc;
};

// An array's tile, element j kBlock + tid in w[j] (0 past n), all loads in
// flight together; the block synchronises, as the log-weight form does.
__device__ __forceinline__ void fe_load_w(WArray s, long long base, long long n,
    double *, double (&w)[kItems])
{
    const bool full = base + kTile <= n;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const long long g = base + j * kBlock + threadIdx.x;
        w[j] = s.w[full || g < n ? g : n - 1];
    }
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        keep_loaded(w[j]);
        const long long g = base + j * kBlock + threadIdx.x;
        w[j] = full || g < n ? w[j] : 0.0;
    }
    __syncthreads();
}

// A tile's weights, element j kBlock + tid in w[j] (0 past n): the
// log-weights are loaded first (every load in flight before the first is
// used; clamped indices in a partial tile), then W from each, with e^x from
// the LDS copy tab of the table, which this copies and synchronises.
__device__ __forceinline__ void fe_load_w(WLogV s, long long base, long long n,
    double *tab, double (&w)[kItems])
{
    double l[kItems];
    const bool full = base + kTile <= n;
    if (full) {
#pragma unroll
        for (int j = 0; j < kItems; ++j)
            l[j] = s.raw(base + j * kBlock + threadIdx.x);
    } else {
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const long long g = base + j * kBlock + threadIdx.x;
            l[j] = s.raw(g < n ? g : n - 1);
        }
    }
    wexp_table_to_lds(tab);
    __syncthreads();
    s.tab = tab;
#pragma unroll
    for (int j = 0; j < kItems; ++j)
        keep_loaded(l[j]);
    if (full) {
#pragma unroll
        for (int j = 0; j < kItems; ++j)
            w[j] = s.from_raw(l[j]);
    } else {
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const long long g = base + j * kBlock + threadIdx.x;
            w[j] = g < n ? s.from_raw(l[j]) : 0.0;
        }
    }
}

// Raw residuals modf(N W_i) (WResidRaw) of a tile from the weights' own
// loader (0 past n).
template <class B>
__device__ __forceinline__ void fe_load_w(WResidRaw<B> s, long long base, long long n,
    double *tab, double (&w)[kItems])
{
    fe_load_w(s.b, base, n, tab, w);
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        double ip;
        w[j] = modf(s.coeff * w[j], &ip);
    }
}

// Normalised residual weights inv_resid * modf(N W_i) (WResidV) of a tile
// from the weights' own loader (0 past n, and when no residual draw
// remains).
template <class B>
__device__ __forceinline__ void fe_load_w(WResidV<B> s, long long base, long long n,
    double *tab, double (&w)[kItems])
{
    fe_load_w(s.b, base, n, tab, w);
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        double ip;
        const double r = modf(s.coeff * w[j], &ip);
        w[j] = s.active ? s.inv_resid * r : 0.0;
    }
}

// Loads a tile's weights and leaves them in blocked order: thread tid holds
// elements tid kItems + j in w[j] (through the LDS tile, element k at
// lds_pad(k), which keeps them afterwards).
template <class V>
__device__ __forceinline__ void fe_load_blocked(V s, long long base,
    long long n, double *tile, double *tab, double (&w)[kItems])
{
    fe_load_w(s, base, n, tab, w);
#pragma unroll
    for (int j = 0; j < kItems; ++j)
        tile[lds_pad(j * kBlock + threadIdx.x)] = w[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kItems; ++j)
        w[j] = tile[lds_pad(threadIdx.x * kItems + j)];
}

// Rounded steps in binade e: q = w / 2^(e - 52) rounded to the nearest
// integer, ties to even, and whether q lies exactly halfway. For 0 <= q <
// 2^52 (every step of a tile that stays in binade e) (q + 2^52) - 2^52 is
// that rounding (the library builds without FP contraction). A larger q
// gives a step of at least 2^52, which fails every check that the sums stay
// below 2^53, so its value does not matter. ok: 2^(52 - e) is a normal
// double (e >= -971; a tile whose sums are smaller takes another path).
struct FeSteps {
    double sc;    // 2^(52 - e)
    double ulp;   // 2^(e - 52)
    int e;
    bool ok;
    __device__ explicit FeSteps(int e_) : e(e_)
    {
        ok = 52 - e_ <= 1023;
        sc = ldexp(1.0, ok ? 52 - e_ : 0);
        ulp = ldexp(1.0, e_ - 52);
    }
    __device__ double q(double w) const { return w * sc; }
    __device__ static double round(double q)
    {
        const double M = 0x1p52;
        return (q + M) - M;
    }
    __device__ static bool half(double q, double r)
    {
        const double d = q - r;
        return d == 0.5 || d == -0.5;
    }
    // a * 2^(e - 52) for an integer a < 2^53 (exact)
    __device__ double value(double a) const { return a * ulp; }
};

struct FeSummShared {
    fold::SegShared seg;         // seg.tile: the tile; the rest: a full summary
    double tab[128];
    double dred[kWaves];
    double dred2[kWaves];
    int ired[kWaves];
    double cross[2];             // the sum after the crossing element, its weight
};

// Tile summaries. Blocks [0, kFePred) take the tiles listed in fpred ([0]
// the count; none when it exceeds kFePred), block kFePred + t takes tile t
// unless a first block took it. n: *ndev when given (the grid covers nh),
// else nh.
template <class WSrc>
__global__ void __launch_bounds__(kBlock) k_fe_summ(WSrc src, long long nh,
    const long long *ndev, const ResampleDev *prm, const double *tpre,
    const int *fpred, TileSeg *segs, FeSum *fs, FeCross *fx)
{
    if (!prm->active)
        return;
    const long long n = ndev ? *ndev : nh;
    const long long ntiles = ceil_div(n, kTile);
    const int npred = fpred[0];
    const bool early = npred <= kFePred;
    long long t;
    if (blockIdx.x < kFePred) {
        if (!early || static_cast<int>(blockIdx.x) >= npred)
            return;
        t = fpred[1 + blockIdx.x];
    } else {
        t = static_cast<long long>(blockIdx.x) - kFePred;
        if (t >= ntiles || (early && fold::predicted_crossing(tpre, t, ntiles)))
            return;
    }
    const long long base = t * kTile;
    __shared__ FeSummShared sh;
    const int tid = static_cast<int>(threadIdx.x);
    const long long rem = n - base;
    const int nvalid = rem < kTile ? static_cast<int>(rem) : kTile;
    double w[kItems];
    fe_load_w(src.loaded(), base, n, sh.tab, w);
    const double x0 = t == 0 ? 0.0 : tpre[t];
    const bool pos = x0 > 0.0 && x0 < INFINITY;
    const FeSteps st(pos ? fold::binade(x0) : 0);
    {
        double wsum = 0.0, dsum = 0.0;
        bool tie = false;
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const double q = st.q(w[j]);
            const double r = FeSteps::round(q);
            tie = tie || FeSteps::half(q, r);
            wsum += w[j];
            dsum += r;
        }
        const double T = block_reduce(wsum, OpAddD(), sh.dred);
        const double D = block_reduce(dsum, OpAddD(), sh.dred2);
        const bool anytie = __syncthreads_or(tie);
        const double end = x0 + T;
        if (pos && st.ok && !anytie && end < INFINITY && fold::binade(end) == st.e) {
            // one binade, no ties: the map adds the integer total of the
            // steps (saturated at 2^54: such a tile certainly leaves it)
            if (tid == 0) {
                const long long d = D < static_cast<double>(fold::kCap) ?
                    static_cast<long long>(D) : fold::kCap;
                segs[t].ncross = 0;
                segs[t].e[0] = st.e;
                segs[t].map[0] = FMap{d, d};
                fs[t] = FeSum{kFeMap, st.e, d, d};
            }
            return;
        }
    }
    // the tile in blocked order (and in LDS, element k at lds_pad(k))
#pragma unroll
    for (int j = 0; j < kItems; ++j)
        sh.seg.tile[lds_pad(j * kBlock + tid)] = w[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kItems; ++j)
        w[j] = sh.seg.tile[lds_pad(tid * kItems + j)];
    if (pos && st.ok) {
        // one crossing of the approximate sums: an element whose sum lies
        // in another binade than the sum before it
        double loc = 0.0;
#pragma unroll
        for (int j = 0; j < kItems; ++j)
            loc += w[j];
        double tw;
        const double ex = block_exclusive_scan(loc, OpAddD(), 0.0, sh.dred, &tw);
        int ncr = 0;
        int kc = kTile;
        double pre = x0 + ex;
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const double post = pre + w[j];
            const bool c = fold::sum_tag(pre) != fold::sum_tag(post);
            ncr += c ? 1 : 0;
            if (c && kc == kTile) {
                kc = tid * kItems + j;
                sh.cross[0] = post;
                sh.cross[1] = w[j];
            }
            pre = post;
        }
        const int nc_tot = block_reduce(ncr, fold::OpAddI(), sh.ired);
        const int c = block_reduce(kc, OpMinI(), sh.ired);
        if (nc_tot == 1 && c < nvalid) {
            const double after = sh.cross[0];
            const double wcr = sh.cross[1];
            const bool aok = after < INFINITY && after > 0.0;
            const FeSteps st2(aok ? fold::binade(after) : 0);
            double d1s = 0.0, d2s = 0.0;
            bool ties = false;
#pragma unroll
            for (int j = 0; j < kItems; ++j) {
                const int k = tid * kItems + j;
                const double q = k < c ? st.q(w[j]) : st2.q(w[j]);
                const double r = FeSteps::round(q);
                ties = ties || (FeSteps::half(q, r) && k != c);
                d1s += k < c ? r : 0.0;
                d2s += k > c ? r : 0.0;
            }
            const double D1 = block_reduce(d1s, OpAddD(), sh.dred);
            const double D2 = block_reduce(d2s, OpAddD(), sh.dred2);
            const bool anyt = __syncthreads_or(ties);
            if (aok && st2.ok && !anyt && D1 < static_cast<double>(fold::kCap) &&
                D2 < static_cast<double>(fold::kCap)) {
                if (tid == 0) {
                    const long long a = static_cast<long long>(D1);
                    segs[t].ncross = -1;
                    fs[t] = FeSum{kFeCross, st.e, a, a};
                    fx[t] = FeCross{st2.e, 0, static_cast<long long>(D2), wcr};
                }
                return;
            }
        }
    }
    // the full summary, from the tile in LDS
    __syncthreads();
    fold::tile_segments(nvalid, x0, &segs[t], sh.seg);
    __syncthreads();
    if (tid == 0) {
        const int nc = segs[t].ncross;
        const int e0 = segs[t].e[0];
        const FMap m = segs[t].map[0];
        fs[t] = (nc == 0 && e0 != fold::kEmpty) ? FeSum{kFeMap, e0, m.d0, m.d1} :
                                                 FeSum{kFeList, fold::kEmpty, 0, 0};
    }
}

// A thread's tiles for the resolver's scan: heads and full summaries
// (packed counts) and the composed map since its last head; combined, the
// map restarts at a right operand that holds a head.
struct FeSegScan {
    int hl;
    FMap m;
};
struct OpFeSegScan {
    __device__ FeSegScan operator()(FeSegScan a, FeSegScan b) const
    {
        return FeSegScan{a.hl + b.hl, (b.hl & 0xFFFF) ? b.m : fold::compose(a.m, b.m)};
    }
};

constexpr int kFeThreads = kBlock;     // resolver block (its fallback is kBlock-wide)
constexpr int kFeWaves = kFeThreads / kWave;

struct FeResolveShared {
    TileSeg seg[kFeSegs];       // full summaries of the chunk's listed tiles
    int segtile[kFeSegs];
    // the chain: segments of one-binade tiles and the other tiles, in order
    int ce[kFeChainMax];        // segment binade; listed: kFeListTag - slot;
                                // one crossing: kFeCrossTag - binade index
    FMap cT[kFeChainMax];       // segment map, or the crossing tile's d0
    FeCross cx2[kFeChainMax];   // a crossing tile's second run
    double cx[kFeChainMax];     // exact value before the item
    FeSegScan pscan[kFeWaves];
    int lastkind[kFeThreads];   // each thread's last tile: its kind, or -1 past the end
    int laste[kFeThreads];
    int firsthead[kFeThreads];  // each thread's first tile starts a chain item
    double carry;
    int fail;
};

union FeResolveLds {
    FeResolveShared f;
    ResolveShared r;
};

// The double a * 2^(e - 52) of an integer 2^52 <= a < 2^53 (or a < 2^53 in
// the lowest binade, e = -1022) from its bits; fold::from_mant otherwise.
__device__ inline double fe_from_mant(long long a, int e)
{
    if (e == -1022)
        return __longlong_as_double(a);
    if (a >= (1LL << 52))
        return __longlong_as_double((static_cast<long long>(e + 1022) << 52) + a);
    return fold::from_mant(a, e);
}

// Advances the exact value x by a run in binade e with map m (false: x is
// not in that binade, or the run leaves it).
__device__ inline bool fe_run(double &x, int e, FMap m)
{
    if (!(x > 0.0) || !(x < INFINITY) || fold::binade(x) != e)
        return false;
    const long long a0 = fold::mant(x);
    const long long a = a0 + ((a0 & 1) ? m.d1 : m.d0);
    if (a >= fold::kTwo53)
        return false;
    x = fe_from_mant(a, e);
    return true;
}

// fold::apply_segments with the runs' data read four at a time (a chain
// step waits on the previous value, not on LDS).
__device__ inline bool fe_apply_segments(const TileSeg &sg, double &x)
{
    const int nc = sg.ncross;
    if (nc < 0)
        return false;
    for (int r0 = 0; r0 <= nc; r0 += 4) {
        int e4[4];
        FMap m4[4];
        double c4[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int r = r0 + k <= fold::kSegMax ? r0 + k : fold::kSegMax;
            e4[k] = sg.e[r];
            m4[k] = sg.map[r];
            c4[k] = sg.cross[r < fold::kSegMax ? r : fold::kSegMax - 1];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int r = r0 + k;
            if (r > nc)
                return true;
            if (e4[k] != fold::kEmpty && !fe_run(x, e4[k], m4[k]))
                return false;
            if (r < nc)
                x = x + c4[k];
        }
    }
    return true;
}

// Exact tile starts from the summaries; see the file comment. One block.
// n as k_fe_summ's.
template <class WSrc>
__global__ void __launch_bounds__(kFeThreads) k_fe_resolve(WSrc src, long long nh,
    const long long *ndev, const ResampleDev *prm, const TileSeg *segs,
    const FeSum *fs, const FeCross *fx, GroupSeg *groups, double *xstart,
    double *xgstart, int *gmode)
{
    __shared__ FeResolveLds u;
    FeResolveShared &sh = u.f;
    const long long n = ndev ? *ndev : nh;
    if (n <= 0)
        return;
    const long long ntiles = ceil_div(n, kTile);
    const int tid = static_cast<int>(threadIdx.x);
    // the first chunk's summaries load with the event flag
    FeSum v[kFePer];
#pragma unroll
    for (int k = 0; k < kFePer; ++k) {
        const long long t = static_cast<long long>(tid) * kFePer + k;
        v[k] = fs[t < ntiles ? t : ntiles - 1];
    }
    if (!prm->active)
        return;
    if (tid == 0) {
        sh.carry = 0.0;
        sh.fail = 0;
    }
    __syncthreads();
    for (long long c0 = 0; c0 < ntiles && !sh.fail; c0 += kFeChunk) {
        const long long tb = c0 + static_cast<long long>(tid) * kFePer;
        if (c0 > 0) {
#pragma unroll
            for (int k = 0; k < kFePer; ++k) {
                const long long t = tb + k;
                v[k] = fs[t < ntiles ? t : ntiles - 1];
            }
        }
        // kind per tile, -1 past the end; a map tile joins the segment of
        // the map tile before it when both lie in one binade
        int kind[kFePer];
        int ee[kFePer];
        FMap mm[kFePer];
#pragma unroll
        for (int k = 0; k < kFePer; ++k) {
            kind[k] = tb + k >= ntiles ? -1 : v[k].kind;
            ee[k] = v[k].e;
            mm[k] = FMap{v[k].d0, v[k].d1};
        }
        sh.lastkind[tid] = kind[kFePer - 1];
        sh.laste[tid] = ee[kFePer - 1];
        __syncthreads();
        int pk = tid == 0 ? kFeList : sh.lastkind[tid - 1];
        int pe = tid == 0 ? 0 : sh.laste[tid - 1];
        unsigned head = 0;
        int nhead = 0, nlist = 0;
        FMap agg = fold::ident();
#pragma unroll
        for (int k = 0; k < kFePer; ++k) {
            bool h = false;
            if (kind[k] == kFeMap)
                h = (tb + k == c0) || pk != kFeMap || pe != ee[k];
            else if (kind[k] >= 0)
                h = true;
            head |= h ? (1u << k) : 0u;
            nhead += h ? 1 : 0;
            nlist += kind[k] == kFeList ? 1 : 0;
            const FMap mk = kind[k] == kFeMap ? mm[k] : fold::ident();
            agg = h ? mk : fold::compose(agg, mk);
            pk = kind[k];
            pe = ee[k];
        }
        // one scan: the segmented maps and both counts (each below 2^15)
        FeSegScan stot;
        const FeSegScan sc = block_exclusive_scan_n<kFeThreads>(
            FeSegScan{nhead | (nlist << 16), agg}, OpFeSegScan(),
            FeSegScan{0, fold::ident()}, sh.pscan, &stot);
        const int hrank = sc.hl & 0xFFFF, lrank = sc.hl >> 16;
        const int tot_head = stot.hl & 0xFFFF, tot_list = stot.hl >> 16;
        if (tot_head > kFeChainMax || tot_list > kFeSegs) {
            if (tid == 0)
                sh.fail = 1;
            __syncthreads();
            break;
        }
        sh.firsthead[tid] = ((head & 1u) != 0 || kind[0] < 0) ? 1 : 0;
        __syncthreads();
        {
            int r = hrank - 1;   // chain item of the current tile
            int li = lrank;
            FMap acc = sc.m;
#pragma unroll
            for (int k = 0; k < kFePer; ++k) {
                if (kind[k] < 0)
                    break;
                const bool h = (head >> k) & 1u;
                const FMap mk = kind[k] == kFeMap ? mm[k] : fold::ident();
                acc = h ? mk : fold::compose(acc, mk);
                if (h) {
                    ++r;
                    if (kind[k] == kFeList) {
                        sh.ce[r] = kFeListTag - li;
                        sh.segtile[li] = static_cast<int>(tb + k);
                        ++li;
                    } else if (kind[k] == kFeCross) {
                        sh.ce[r] = kFeCrossTag - (ee[k] + 1100);
                        sh.cT[r] = mm[k];
                        sh.cx2[r] = fx[tb + k];
                    } else {
                        sh.ce[r] = ee[k];
                    }
                }
                bool next_head;
                if (k + 1 < kFePer)
                    next_head = ((head >> (k + 1)) & 1u) != 0 || kind[k + 1] < 0;
                else
                    next_head = tid + 1 == kFeThreads || sh.firsthead[tid + 1] != 0;
                if (kind[k] == kFeMap && next_head)
                    sh.cT[r] = acc;
            }
        }
        __syncthreads();
        // full summaries of the listed tiles into LDS
        constexpr int kWords = static_cast<int>(sizeof(TileSeg) / 4);
        for (int q = tid; q < tot_list * kWords; q += kFeThreads) {
            const int k = q / kWords, wd = q % kWords;
            reinterpret_cast<int *>(&sh.seg[k])[wd] =
                reinterpret_cast<const int *>(&segs[sh.segtile[k]])[wd];
        }
        __syncthreads();
        if (tid == 0) {
            // the last item of the data: only its start is needed (the
            // counting pass recomputes the last tile from it, and each
            // tile's start is checked below), so its prediction past its
            // start is not verified (the total lies at 1, a binade boundary,
            // where predictions often fail)
            const bool last_chunk = c0 + kFeChunk >= ntiles;
            double x = sh.carry;
            int ok = 1;
            // each item's data is read one item ahead, so that the walk waits
            // on its own chain, not on LDS (an item reads only the fields its
            // kind set; the others are in-bounds values it ignores)
            int ce_n = sh.ce[0];
            FMap T_n = sh.cT[0];
            FeCross c2_n = sh.cx2[0];
            for (int r = 0; r < tot_head; ++r) {
                const bool final_item = last_chunk && r == tot_head - 1;
                const int ce = ce_n;
                const FMap T = T_n;
                const FeCross c2 = c2_n;
                if (r + 1 < tot_head) {
                    ce_n = sh.ce[r + 1];
                    T_n = sh.cT[r + 1];
                    c2_n = sh.cx2[r + 1];
                }
                sh.cx[r] = x;
                bool good;
                if (ce > kFeListTag) {
                    good = fe_run(x, ce, T);
                } else if (ce > kFeCrossTag) {
                    good = fe_apply_segments(sh.seg[kFeListTag - ce], x);
                } else {
                    // one crossing: a run, the real add, a run
                    good = fe_run(x, kFeCrossTag - ce - 1100, T);
                    if (good) {
                        x = x + c2.wc;
                        good = fe_run(x, c2.e2, FMap{c2.d2, c2.d2});
                    }
                }
                if (!good) {
                    if (!final_item)
                        ok = 0;
                    break;
                }
            }
            sh.carry = x;
            sh.fail = ok ? 0 : 1;
        }
        __syncthreads();
        if (sh.fail)
            break;
        // every tile's exact start
        {
            int r = hrank - 1;
            FMap P = fold::ident();   // map from the segment start to this tile
            double xs = 0.0;
            long long a0 = 0;
            int e = 0;
            bool bad = false;
            if (!((head & 1u) != 0) && kind[0] == kFeMap) {
                // continues the previous thread's segment
                P = sc.m;
                xs = sh.cx[r];
                a0 = fold::mant(xs);
                e = ee[0];
            }
#pragma unroll
            for (int k = 0; k < kFePer; ++k) {
                if (kind[k] < 0)
                    break;
                const long long t = tb + k;
                const bool h = (head >> k) & 1u;
                if (h) {
                    ++r;
                    P = fold::ident();
                    xs = sh.cx[r];
                    a0 = fold::mant(xs);
                    e = ee[k];
                }
                if (kind[k] != kFeMap) {
                    xstart[t] = xs;
                } else {
                    const long long a = a0 + ((a0 & 1) ? P.d1 : P.d0);
                    // a start past its binade (only possible in the last
                    // segment, whose end is not checked)
                    bad = bad || a >= fold::kTwo53;
                    xstart[t] = h ? xs : fe_from_mant(a, e);
                    P = fold::compose(P, mm[k]);
                }
            }
            if (__syncthreads_or(bad) && tid == 0)
                sh.fail = 1;
        }
        __syncthreads();
    }
    if (sh.fail) {
        // the general resolver: group summaries, the windowed walk, the
        // starts of the groups advanced whole
        __syncthreads();
        const long long ngroups = ceil_div(ntiles, kGroup);
        for (long long g = threadIdx.x / kWave; g < ngroups; g += kFeWaves)
            group_summary(segs, ntiles, g, groups);
        __syncthreads();
        resolve_body(src, n, segs, groups, xstart, xgstart, gmode, u.r);
        __syncthreads();
        for (long long g = threadIdx.x / kWave; g < ngroups; g += kFeWaves)
            starts_group(g, ntiles, segs, xgstart, gmode, xstart);
    }
}

// Systematic draws with the parameters in registers: u_j = u0 + j delta
// without FMA (u01_sequence.hpp:246, as USys). Counts are below 2^31
// (particle indices are int32).
struct FeSys {
    int n;
    double u0;
    double delta;
    __device__ explicit FeSys(const USys &us)
        : n(static_cast<int>(us.p->ndraw)), u0(us.p->u0), delta(us.p->delta) {}
    __device__ double ui(int j) const
    {
        return rn_add(u0, rn_mul(static_cast<double>(j), delta));
    }
    // Number of draws u_j <= x. The estimate (x - u0) n is within n 2^-50 of
    // its real value and every u_j within 2^-51 of u0 + j / n (x <= 2), so a
    // fractional part farther than n 2^-48 from an integer fixes the count;
    // otherwise the draws next to the estimate are compared (USys::count).
    __device__ int count(double x) const
    {
        const double nd = static_cast<double>(n);
        const double est = (x - u0) * nd;
        const double fl = floor(est);
        const double fr = est - fl;
        const double m = nd * 0x1p-48;
        if (x <= 2.0 && est >= 0.0 && est < nd && fr > m && fr < 1.0 - m)
            return static_cast<int>(fl) + 1;
        return slow(x, est);
    }
    __device__ int slow(double x, double est) const
    {
        if (n <= 0)
            return 0;
        if (!(x < INFINITY))
            return n;
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

// Stratified draws: UStrat's search (two Philox blocks per count).
struct FeStrat {
    UStrat us;
    __device__ explicit FeStrat(const UStrat &u) : us(u) {}
    __device__ int count(double x) const { return static_cast<int>(us.count(x)); }
};

template <class W> struct FeIsResid : std::false_type {};
template <class B> struct FeIsResid<WResid<B>> : std::true_type {};

template <class U> struct FeCounter;
template <> struct FeCounter<USys> { using type = FeSys; };
template <> struct FeCounter<UStrat> { using type = FeStrat; };

// Counts of a tile's exact sums (thread tid holds those of elements
// k0 = tid kItems .. k0 + kItems - 1 in sk, blocked) against materialised
// sorted draws (multinomial): the tile's counts lie in [glo, ghi], the
// counts at its exact start and at the next tile's (k_fe_arr_bounds), so
// only the slice v[glo, ghi) decides them. The slice goes through LDS (buf,
// the tile buffer) in chunks of kFeSlice draws (a tile's share of the draws
// and some margin), every open sum searched there (a binary search for a
// thread's first, then rounds of kFeProbes probes past the previous answer,
// since a thread's sums are sorted); a
// slice of more than kVChunks chunks (weight concentrated in the tile) is
// searched in place by every sum, so that no block walks a long slice
// alone. g[j] = G(sk[j]) and gp = G of the sum before the thread's first
// element (glo for the first thread).
// The draws of a count kernel's slice: materialised sorted uniforms
// (multinomial; loaded) or stratified draws (generated from their Philox
// words, v_j = max(u_{j-1}, u_j) as UStrat). at(j) is draw j; fill(buf, j0,
// cn) puts draws j0 .. j0 + cn - 1 at buf[0 ..] and +inf in the chunk's
// other slots (block-wide; the caller synchronises after it).
template <class U> struct FeDraws;

template <> struct FeDraws<UArr> {
    const double *v;
    __device__ explicit FeDraws(const UArr &us) : v(us.v) {}
    __device__ double at(long long j) const { return v[j]; }
    __device__ void fill(double *buf, long long j0, int cn) const;
};

template <> struct FeDraws<UStrat> {
    UStrat us;
    __device__ explicit FeDraws(const UStrat &u) : us(u) {}
    __device__ double at(long long j) const { return us.vi(static_cast<int>(j)); }
    __device__ void fill(double *buf, long long j0, int cn) const;
};

constexpr int kFeProbes = 8;                        // draws probed at once
constexpr int kFeSliceItems = kTileLds / kBlock;     // chunk slots per thread
constexpr int kFeSlice = kFeSliceItems * kBlock - kFeProbes;   // draws per chunk
static_assert(kFeSliceItems * kBlock <= kTileLds, "a chunk fits the tile buffer");

// clamped, unconditional loads, all in flight together
__device__ inline void FeDraws<UArr>::fill(double *buf, long long j0, int cn) const
{
    const int tid = static_cast<int>(threadIdx.x);
    const double *const vc = v + j0;
    double x[kFeSliceItems];
#pragma unroll
    for (int j = 0; j < kFeSliceItems; ++j) {
        const int q = j * kBlock + tid;
        x[j] = vc[q < cn ? q : cn - 1];
    }
#pragma unroll
    for (int j = 0; j < kFeSliceItems; ++j) {
        keep_loaded(x[j]);
        buf[j * kBlock + tid] = j * kBlock + tid < cn ? x[j] : INFINITY;
    }
}

// u_w for the stream words of draws j0 - 1 .. j0 + cn - 1 (a Philox block,
// four words, per thread at a time, from the block holding draw j0 - 1's
// word), then v_j = max(u_{j-1}, u_j) (v_0 = u_0) in place.
__device__ inline void FeDraws<UStrat>::fill(double *buf, long long j0, int cn) const
{
    const int tid = static_cast<int>(threadIdx.x);
    const ResampleDev *const p = us.p;
    const unsigned int key = p->key;
    const double delta = p->delta;
    const long long jf = j0 > 0 ? j0 - 1 : 0;        // first draw generated
    const unsigned long long wf = p->word0 + static_cast<unsigned long long>(jf);
    const unsigned long long b0 = wf / 4;             // its Philox block - 1
    const long long lead = static_cast<long long>(wf - 4 * b0);   // slot of draw jf
    // slots 0 .. nslot - 1 used (nslot <= 4 + kFeSlice: 32-bit slot indices;
    // draw jf + s - lead as jf + (s - lead) in FP64, both exact integers)
    const int nslot = static_cast<int>(lead + (j0 + cn - jf));
    const int nblk = (nslot + 3) / 4;
    const int ilead = static_cast<int>(lead);
    const double djf = static_cast<double>(jf);
    for (int b = tid; b < nblk; b += kBlock) {
        const u32x4 w = philox_block(key, b0 + static_cast<unsigned long long>(b) + 1);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int slot = 4 * b + q;
            const double j = djf + static_cast<double>(slot - ilead);   // draw of this word
            if (slot >= ilead && slot < nslot)
                buf[slot] = rn_add(rn_mul(u01_u32d(w.v[q]), delta), rn_mul(j, delta));
        }
    }
    __syncthreads();
    // draw j0 + q at slot lead + (j0 - jf) + q; v from the slot before. In
    // groups of kG items, each group's reads before its writes, so that only
    // kG values are live (the caller's sums stay in registers). A group's
    // writes reach the slots the next group reads only at the slot before
    // its first draw, and only when s0 == 0: that raw value is kept in
    // edge[] (one per group parity: a group reads one while it writes the
    // other).
    const long long s0 = lead + (j0 - jf);
    constexpr int kG = 6;
    __shared__ double edge[2];
#pragma unroll
    for (int g0 = 0; g0 < kFeSliceItems; g0 += kG) {
        const int par = (g0 / kG) & 1;
        double x[kG];
#pragma unroll
        for (int jj = 0; jj < kG; ++jj) {
            const int j = g0 + jj;
            if (j >= kFeSliceItems)
                break;
            const int q = j * kBlock + tid;
            double a = 0.0;
            if (q < cn) {
                a = buf[s0 + q];
                if ((jj == kG - 1 || j == kFeSliceItems - 1) && tid == kBlock - 1)
                    edge[par] = a;
                if (j0 + q > 0) {
                    const double b = (g0 > 0 && jj == 0 && tid == 0 && s0 == 0)
                        ? edge[par ^ 1] : buf[s0 + q - 1];
                    a = b > a ? b : a;
                }
            }
            x[jj] = a;
        }
        __syncthreads();
#pragma unroll
        for (int jj = 0; jj < kG; ++jj) {
            const int j = g0 + jj;
            if (j >= kFeSliceItems)
                break;
            const int q = j * kBlock + tid;
            buf[q] = q < cn ? x[jj] : INFINITY;
        }
        if (g0 + kG < kFeSliceItems)
            __syncthreads();
    }
}

template <class U>
__device__ inline void fe_arr_counts(const FeDraws<U> &us, const double (&sk)[kItems],
    long long glo, long long ghi, int nvalid, double *buf, int (&g)[kItems],
    int &gp)
{
    const int tid = static_cast<int>(threadIdx.x);
    const int k0 = tid * kItems;
    const long long m = ghi > glo ? ghi - glo : 0;
    const bool in_lds = m <= static_cast<long long>(kVChunks) * kFeSlice;
    int cnt[kItems];   // counts within the slice (m < 2^31)
    unsigned open = 0;   // bit j: the count of sum j not found yet
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        cnt[j] = static_cast<int>(m);
        open |= k0 + j < nvalid ? 1u << j : 0u;
    }
    if (!in_lds) {
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            if ((open >> j) & 1u) {
                long long lo = 0, hi = m;   // first slice index with v > sk
                while (lo < hi) {
                    const long long mid = (lo + hi) >> 1;
                    if (us.at(glo + mid) <= sk[j])
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                cnt[j] = static_cast<int>(lo);
            }
        }
        open = 0;
    }
    for (long long c0 = 0; c0 < (in_lds ? m : 0); c0 += kFeSlice) {
        const int cn = m - c0 < kFeSlice ? static_cast<int>(m - c0) : kFeSlice;
        us.fill(buf, glo + c0, cn);
        __syncthreads();
        // a thread's sums are sorted: a binary search for the first one
        // resolved in this chunk, then from the previous answer a (the
        // first draw above the previous sum) kFeProbes probes at once, about
        // a draw per sum on average (a binary search when all lie below)
        const double vlast = buf[cn - 1];
        auto search = [&](int lo, double x) {
            // first index with buf > x in [lo, cn); buf[cn - 1] > x
            int hi = cn - 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (buf[mid] <= x)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            return lo;
        };
        int a = -1;
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            if (((open >> j) & 1u) && sk[j] < vlast) {
                if (a < 0) {
                    a = search(0, sk[j]);
                } else {
                    // a < cn, so the probes stay in the buffer (slots past
                    // the chunk: +inf); the draws at or below sk[j] are a
                    // prefix of them. A second round when all lie below
                    // (then a + kFeProbes <= the answer < cn), a binary
                    // search after that (a wave runs a round when any lane
                    // needs it)
                    auto probe = [&](int at) {
                        int c = 0;
#pragma unroll
                        for (int q = 0; q < kFeProbes; ++q)
                            c += buf[at + q] <= sk[j] ? 1 : 0;
                        return c;
                    };
                    int n = probe(a);
                    a += n;
                    if (n == kFeProbes) {
                        n = probe(a);
                        a += n;
                        if (n == kFeProbes)
                            a = search(a, sk[j]);
                    }
                }
                cnt[j] = static_cast<int>(c0) + a;
                open &= ~(1u << j);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < kItems; ++j)
        g[j] = static_cast<int>(glo + cnt[j]);
    // G before the thread's first element: the previous thread's last
    int *const last = reinterpret_cast<int *>(buf);
    last[tid] = g[kItems - 1];
    __syncthreads();
    gp = tid == 0 ? static_cast<int>(glo) : last[tid - 1];
}

// The draw counts G(x) = #{j : v_j <= x} at every tile's exact start x =
// xstart[t] into gb[t] (0 for tile 0), one wave per tile, 64 probes a
// round; the counting pass reads its slice bounds there.
__global__ void __launch_bounds__(kBlock) k_fe_arr_bounds(UArr us, long long M,
    const ResampleDev *prm, const double *xstart, int *gb)
{
    if (!prm->active)
        return;
    const long long ntiles = (M + kTile - 1) / kTile;
    const long long t = static_cast<long long>(blockIdx.x) * kWaves +
        static_cast<long long>(threadIdx.x) / kWave;
    if (t >= ntiles)
        return;
    const long long g = t == 0 ? 0 :
        wave_last_le([&](long long q) { return us.v[q]; }, us.ndraw(), xstart[t]) + 1;
    if ((threadIdx.x & (kWave - 1)) == 0)
        gb[t] = static_cast<int>(g);
}

// The same for stratified draws, by UStrat's own search: one thread per
// tile.
__global__ void __launch_bounds__(kBlock) k_fe_strat_bounds(UStrat us, long long M,
    const ResampleDev *prm, const double *xstart, int *gb)
{
    if (!prm->active)
        return;
    const long long ntiles = (M + kTile - 1) / kTile;
    const long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= ntiles)
        return;
    gb[t] = t == 0 ? 0 : static_cast<int>(us.count(xstart[t]));
}

// Relative bound between an FP64 block-scan prefix of a tile (exact start
// plus a blocked scan: at most 16 + 6 + 3 + 18 roundings along any path)
// and the sequential fold over at most 4096 weights from the same exact
// start: (4097 + 43) 2^-53 (1 + 2^-40) < 2^-40.9; twice that, so that the
// rounding of the bounds themselves is covered.
constexpr double kFeCertRel = 0x1p-39;

struct FeCountShared {
    double tile[kTileLds];
    fold::ExactShared ex;
    double tab[128];
    double dlds[kWaves];
    long long lscan[kWaves];
};

// Counts of every tile from its exact start xstart[t] (see the file
// comment); rep -> index aggregates (childless, extras, sources) per tile
// and the tile's sources (csrc, cE) in its own region.
template <class WSrc, class USrc>
__global__ void __launch_bounds__(kBlock, 4) k_fe_count(WSrc src, USrc us,
    long long M, ResampleDev *prm, double *xstart, int *rep,
    long long *tagg, int *csrc, int *cE, const int *gb)
{
    if (!prm->active)
        return;
    const long long t = blockIdx.x;
    const long long base = t * kTile;
    if (base >= M)
        return;
    __shared__ FeCountShared sh;
    const int tid = static_cast<int>(threadIdx.x);
    const long long rem = M - base;
    const int nvalid = rem < kTile ? static_cast<int>(rem) : kTile;
    const bool last_tile = base + kTile >= M;
    const double x0 = xstart[t];
    // sorted or stratified draws: the tile's slice bounds (the last tile's
    // reaches the end)
    constexpr bool kSlice = std::is_same<USrc, UArr>::value ||
        std::is_same<USrc, UStrat>::value;
    long long glo = 0, ghi = 0;
    if constexpr (kSlice) {
        glo = gb[t];
        ghi = last_tile ? us.ndraw() : gb[t + 1];
    }
    double w[kItems];
    fe_load_blocked(src.loaded(), base, M, sh.tile, sh.tab, w);
    const bool xpos = x0 > 0.0 && x0 < INFINITY;
    const FeSteps st(xpos ? fold::binade(x0) : 0);
    double dsum = 0.0;
    bool tie = false;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const double q = st.q(w[j]);
        const double r = FeSteps::round(q);
        tie = tie || FeSteps::half(q, r);
        dsum += r;
    }
    double D;
    double run = block_exclusive_scan(dsum, OpAddD(), 0.0, sh.dlds, &D);
    const bool anytie = __syncthreads_or(tie);
    const double a0d = xpos ? static_cast<double>(fold::mant(x0)) : 0.0;
    const bool fast = xpos && st.ok && !anytie &&
        a0d + D < static_cast<double>(fold::kTwo53);
    const int nd = static_cast<int>(us.ndraw());
    const int k0 = tid * kItems;
    // g[j] = G(s_k) for element k = k0 + j, gp = G before the thread's first
    int g[kItems];
    int gp;
    if constexpr (kSlice) {
        // sorted or stratified draws: the exact sums first (from the steps,
        // or the exact fold), then their counts from the tile's slice
        double sk[kItems];
        if (fast) {
#pragma unroll
            for (int j = 0; j < kItems; ++j) {
                run += FeSteps::round(st.q(w[j]));
                sk[j] = st.value(a0d + run);
            }
        } else {
            fold::tile_exact(sh.tile, x0, sh.ex);
#pragma unroll
            for (int j = 0; j < kItems; ++j)
                sk[j] = sh.tile[lds_pad(k0 + j)];
        }
        __syncthreads();   // every thread's sums read from the tile
        fe_arr_counts(FeDraws<USrc>(us), sk, glo, ghi, nvalid, sh.tile, g, gp);
    } else {
        const typename FeCounter<USrc>::type cnt(us);
        if (fast) {
            // s_k = x0 advanced by the steps up to k, all in x0's binade
            const double sprev = tid == 0 ? x0 : st.value(a0d + run);
            gp = (base + k0 == 0) ? 0 : cnt.count(sprev);
            double s[kItems];
#pragma unroll
            for (int j = 0; j < kItems; ++j) {
                run += FeSteps::round(st.q(w[j]));
                s[j] = st.value(a0d + run);
                g[j] = cnt.count(s[j]);
            }
            // the total after the last tile, which the resolver leaves
            // unverified (the fast event's callers do not read it)
            if (last_tile) {
#pragma unroll
                for (int j = 0; j < kItems; ++j)
                    if (k0 + j == nvalid - 1)
                        xstart[t + 1] = s[j];
            }
        } else {
            // an FP64 prefix within kFeCertRel of every exact sum: a count is
            // certain when both ends of the interval give it
            double loc = 0.0;
#pragma unroll
            for (int j = 0; j < kItems; ++j)
                loc += w[j];
            double tw;
            const double ex = block_exclusive_scan(loc, OpAddD(), 0.0, sh.dlds, &tw);
            double p = x0 + ex;
            bool amb = false;
            auto certain = [&](double v) {
                if (!(v < INFINITY))
                    return cnt.count(v);
                const double m = v * kFeCertRel;
                const int lo = cnt.count(v - m);
                const int hi = cnt.count(v + m);
                amb = amb || lo != hi;
                return lo;
            };
            if (base + k0 == 0)
                gp = 0;
            else if (tid == 0)
                gp = cnt.count(x0);
            else
                gp = certain(p);
#pragma unroll
            for (int j = 0; j < kItems; ++j) {
                p = p + w[j];
                g[j] = certain(p);
            }
            if (__syncthreads_or(amb)) {
                // a draw within the bound of a sum: the exact fold of fold.hpp
                // (one binade with ties in order, runs between crossings, and
                // element by element where a run's prediction fails)
                fold::tile_exact(sh.tile, x0, sh.ex);
                gp = (base + k0 == 0) ? 0 :
                    cnt.count(tid == 0 ? x0 : sh.tile[lds_pad(k0 - 1)]);
#pragma unroll
                for (int j = 0; j < kItems; ++j)
                    g[j] = cnt.count(sh.tile[lds_pad(k0 + j)]);
            }
        }
    }
    // counts: rep_k = G(s_k) - G(s_{k-1}); G before particle 0 is 0, and the
    // last particle takes every remaining draw (residual schemes add the
    // integer parts below)
    int r[kItems];
    if (!last_tile) {
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            r[j] = g[j] - gp;
            gp = g[j];
        }
    } else {
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const int k = k0 + j;
            int gk = gp;
            if (k < nvalid)
                gk = (base + k == M - 1) ? nd : g[j];
            r[j] = gk - gp;
            // words of the stream the event consumed (transform.hpp:61-69:
            // one per draw up to the first failing comparison; systematic:
            // the constructor's one), from the draws of the last particle
            if (base + k == M - 1) {
                const long long jend = gp;
                prm->words_used = std::is_same<USrc, USys>::value ? 1ULL :
                    static_cast<unsigned long long>(nd < 1 ? 0 :
                        (jend + 1 < nd ? jend + 1 : nd));
            }
            gp = gk;
        }
    }
    if constexpr (FeIsResid<WSrc>::value) {
        // residual schemes: rep_k adds floor(N W_k) (transform.hpp:159-161),
        // from the log-weights of the thread's elements again (all loads in
        // flight together; the table copy in LDS is still in place)
        WLogV b = src.b.loaded();
        b.tab = sh.tab;
        double lr[kItems];
#pragma unroll
        for (int j = 0; j < kItems; ++j)
            lr[j] = b.raw(base + (k0 + j < nvalid ? k0 + j : nvalid - 1));
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            keep_loaded(lr[j]);
            double ip;
            modf(src.coeff * b.from_raw(lr[j]), &ip);
            if (k0 + j < nvalid)
                r[j] += static_cast<int>(ip);
        }
    }
    if (k0 + kItems <= nvalid && (reinterpret_cast<uintptr_t>(rep) & 15) == 0) {
        int4 *dst = reinterpret_cast<int4 *>(rep + base + k0);
#pragma unroll
        for (int q = 0; q < kItems / 4; ++q)
            dst[q] = make_int4(r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]);
    } else {
#pragma unroll
        for (int j = 0; j < kItems; ++j)
            if (k0 + j < nvalid)
                rep[base + k0 + j] = r[j];
    }
    // rep -> index: the tile's childless slots, extra copies and sources
    // (per-tile aggregates), and its sources with their exclusive prefix of
    // extra copies within the tile, in the tile's region of csrc / cE
    // (elements tidx kTile ..), staged in LDS so the writes are coalesced
    long long loc = 0;   // extras << 32 | childless << 16 | sources
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        if (k0 + j < nvalid) {
            const int rr = r[j];
            loc += rr == 0 ? (1LL << 16) : 0;
            loc += rr > 1 ? ((static_cast<long long>(rr - 1) << 32) | 1) : 0;
        }
    }
    long long tot;
    const long long ex = block_exclusive_scan(loc, OpAddL(), 0LL, sh.lscan, &tot);
    const int nsrc = static_cast<int>(tot & 0xFFFF);
    int *const ls = reinterpret_cast<int *>(sh.tile);        // sources
    int *const le = ls + kTile;                              // extras before each
    {
        int cr = static_cast<int>(ex & 0xFFFF);
        int er = static_cast<int>(ex >> 32);
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            if (k0 + j < nvalid && r[j] > 1) {
                ls[cr] = static_cast<int>(base + k0 + j);
                le[cr] = er;
                ++cr;
                er += r[j] - 1;
            }
        }
    }
    __syncthreads();
    for (int q = tid; q < nsrc; q += kBlock) {
        csrc[base + q] = ls[q];
        cE[base + q] = le[q];
    }
    if (tid == 0) {
        tagg[3 * t + 0] = (tot >> 16) & 0xFFFF;
        tagg[3 * t + 1] = tot >> 32;
        tagg[3 * t + 2] = nsrc;
    }
}

// ------------------------------------------------ sorted draws (multinomial)
//
// U01SequenceSorted (u01_sequence.hpp:141-154): lmax_k = lmax_{k-1} +
// log(U_k) / (n - k), u_k = 1 - exp(lmax_k). With the negated terms
// a_k = -(log(U_k) / (n - k)) >= 0 the sums s_k = -lmax_k are a sequential
// FP64 fold of non-negative terms (rounded the same way by symmetry), so
// the fast event's summaries and resolver give every tile's exact start:
// k_fe_sorted_terms materialises the terms once (with per-tile sums),
// k_fe_sorted_scan the approximate tile prefix and the tiles predicted to
// cross a binade, k_fe_summ / k_fe_resolve the exact starts, and
// k_fe_sorted_prefix every u_k from its exact sum (the steps in the start's
// binade, else the exact fold), recording whether u ever decreases: the
// running maximum (transform.hpp's merge) is then the identity, and
// k_fe_sorted_max applies it only otherwise.

// Terms a_k for k in [0, n) of the stream (key, word0 + k) (n: *ndev when
// given, the grid covering nh; as in the next two), four words of a
// Philox block per thread at a time (a thread's run of four may straddle
// two blocks when word0 is not a multiple of four), and the tile sums.
__global__ void __launch_bounds__(kBlock) k_fe_sorted_terms(long long nh,
    const long long *ndev, ResampleDev *prm, double *terms, double *tsum)
{
    if (!prm->active)
        return;
    if (blockIdx.x == 0 && threadIdx.x == 0)
        prm->nonmono = 0;
    const long long n = ndev ? *ndev : nh;
    const long long base = static_cast<long long>(blockIdx.x) * kTile;
    if (base >= n)
        return;
    const unsigned int key = prm->key;
    const unsigned long long word0 = prm->word0;
    const double nd = static_cast<double>(n);
    __shared__ double lds[kWaves];
    double acc = 0.0;
#pragma unroll
    for (int g = 0; g < kItems / 4; ++g) {
        const long long k0 = base + 4 * (static_cast<long long>(g) * kBlock + threadIdx.x);
        if (k0 >= n)
            continue;
        const unsigned long long w0 = word0 + static_cast<unsigned long long>(k0);
        const unsigned int off = static_cast<unsigned int>(w0 & 3);
        const u32x4 b0 = philox_block(key, w0 / 4 + 1);
        u32x4 b1 = b0;
        if (off != 0)
            b1 = philox_block(key, w0 / 4 + 2);
        // words off .. off + 3 of the two blocks (off is the same for every
        // thread: constant register indices in each case)
        uint32_t wq[4];
        switch (off) {
        case 0: wq[0] = b0.v[0]; wq[1] = b0.v[1]; wq[2] = b0.v[2]; wq[3] = b0.v[3]; break;
        case 1: wq[0] = b0.v[1]; wq[1] = b0.v[2]; wq[2] = b0.v[3]; wq[3] = b1.v[0]; break;
        case 2: wq[0] = b0.v[2]; wq[1] = b0.v[3]; wq[2] = b1.v[0]; wq[3] = b1.v[1]; break;
        default: wq[0] = b0.v[3]; wq[1] = b1.v[0]; wq[2] = b1.v[1]; wq[3] = b1.v[2]; break;
        }
        double a[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double u = u01_u32d(wq[q]);
            const long long k = k0 + q;
            a[q] = -(log(u) / (nd - static_cast<double>(k)));
        }
        if (k0 + 4 <= n) {
            typedef double D2 __attribute__((ext_vector_type(2)));
            D2 *const dst = reinterpret_cast<D2 *>(terms + k0);
            dst[0] = D2{a[0], a[1]};
            dst[1] = D2{a[2], a[3]};
            acc += (a[0] + a[1]) + (a[2] + a[3]);
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (k0 + q < n) {
                    terms[k0 + q] = a[q];
                    acc += a[q];
                }
            }
        }
    }
    acc = block_reduce(acc, OpAddD(), lds);
    if (threadIdx.x == 0)
        tsum[blockIdx.x] = acc;
}

// The approximate exclusive tile prefix of the terms, in place over the
// tile sums, and the tiles predicted to cross a binade (fpred as the
// finalize writes it for the weights). One block of kFeScanThreads.
constexpr int kFeScanThreads = 1024;
__global__ void __launch_bounds__(kFeScanThreads) k_fe_sorted_scan(long long nh,
    const long long *ndev, const ResampleDev *prm, double *tpre, int *fpred)
{
    if (!prm->active)
        return;
    const long long n = ndev ? *ndev : nh;
    const long long nb = (n + kTile - 1) / kTile;
    __shared__ double lds[kFeScanThreads / kWave];
    __shared__ int cnt;
    block_scan_chunked_n<kFeScanThreads, 4>(nb,
        [&](long long b) { return tpre[b]; },
        [&](long long b, double ex) { tpre[b] = ex; }, OpAddD(), 0.0, lds);
    list_predicted(tpre, nb, fpred, &cnt);
}

struct FeSortedShared {
    double tile[kTileLds];
    fold::ExactShared ex;
    double dlds[kWaves];
};

// Every u_k = 1 - exp(-s_k) (u01_sequence.hpp:151-152) from the exact sums
// of its tile (xstart[t] the exact start), written to out; sets
// prm->nonmono when some u_k < u_{k-1}.
__global__ void __launch_bounds__(kBlock) k_fe_sorted_prefix(long long nh,
    const long long *ndev, ResampleDev *prm, const double *terms,
    const double *xstart, double *out, double *tmax)
{
    if (!prm->active)
        return;
    const long long n = ndev ? *ndev : nh;
    const long long t = blockIdx.x;
    const long long base = t * kTile;
    if (base >= n)
        return;
    __shared__ FeSortedShared sh;
    const int tid = static_cast<int>(threadIdx.x);
    const long long rem = n - base;
    const int nvalid = rem < kTile ? static_cast<int>(rem) : kTile;
    const double x0 = xstart[t];
    double w[kItems];
    {
        // the tile in blocked order (and in LDS, element k at lds_pad(k))
        fe_load_w(WArray{terms}, base, n, nullptr, w);
#pragma unroll
        for (int j = 0; j < kItems; ++j)
            sh.tile[lds_pad(j * kBlock + tid)] = w[j];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kItems; ++j)
            w[j] = sh.tile[lds_pad(tid * kItems + j)];
    }
    const bool xpos = x0 > 0.0 && x0 < INFINITY;
    const FeSteps st(xpos ? fold::binade(x0) : 0);
    double dsum = 0.0;
    bool tie = false;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const double q = st.q(w[j]);
        const double r = FeSteps::round(q);
        tie = tie || FeSteps::half(q, r);
        dsum += r;
    }
    double D;
    double run = block_exclusive_scan(dsum, OpAddD(), 0.0, sh.dlds, &D);
    const bool anytie = __syncthreads_or(tie);
    const double a0d = xpos ? static_cast<double>(fold::mant(x0)) : 0.0;
    const bool fast = xpos && st.ok && !anytie &&
        a0d + D < static_cast<double>(fold::kTwo53);
    const int k0 = tid * kItems;
    double u[kItems];
    if (fast) {
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            run += FeSteps::round(st.q(w[j]));
            u[j] = 1.0 - exp(-st.value(a0d + run));
        }
    } else {
        fold::tile_exact(sh.tile, x0, sh.ex);
#pragma unroll
        for (int j = 0; j < kItems; ++j)
            u[j] = 1.0 - exp(-sh.tile[lds_pad(k0 + j)]);
    }
    // u before the thread's first element: the previous thread's last, or
    // for the first thread the previous tile's last (from its exact sum);
    // the tile buffer holds each thread's last once every sum is read
    __syncthreads();
    double *const last = sh.tile;
    last[tid] = u[kItems - 1];
    __syncthreads();
    double prev = tid > 0 ? last[tid - 1] : (t > 0 ? 1.0 - exp(-x0) : 0.0);
    bool dec = false;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        if (k0 + j < nvalid) {
            dec = dec || u[j] < prev;
            prev = u[j];
        }
    }
    if (k0 + kItems <= nvalid) {
        typedef double D2 __attribute__((ext_vector_type(2)));
        D2 *const dst = reinterpret_cast<D2 *>(out + base + k0);
#pragma unroll
        for (int q = 0; q < kItems / 2; ++q)
            dst[q] = D2{u[2 * q], u[2 * q + 1]};
    } else {
#pragma unroll
        for (int j = 0; j < kItems; ++j)
            if (k0 + j < nvalid)
                out[base + k0 + j] = u[j];
    }
    // the tile's maximum, for the running maximum should u decrease
    double mx = -INFINITY;
#pragma unroll
    for (int j = 0; j < kItems; ++j)
        if (k0 + j < nvalid)
            mx = u[j] > mx ? u[j] : mx;
    mx = block_reduce(mx, OpMaxD(), sh.dlds);
    if (__syncthreads_or(dec) && tid == 0)
        atomicOr(&prm->nonmono, 1);
    if (tid == 0)
        tmax[t] = mx;
}

// The running maximum v_k = max(u_0 .. u_k) in place (transform.hpp's merge
// reads the draws through it), only when k_fe_sorted_prefix saw u decrease:
// tile t's carry is the maximum of the earlier tiles' maxima tmax.
__global__ void __launch_bounds__(kBlock) k_fe_sorted_max(long long nh,
    const long long *ndev, const ResampleDev *prm, const double *tmax,
    double *u)
{
    if (!prm->active || !prm->nonmono)
        return;
    const long long n = ndev ? *ndev : nh;
    const long long t = blockIdx.x;
    const long long base = t * kTile;
    if (base >= n)
        return;
    __shared__ double tile[kTileLds];
    __shared__ double lds[kWaves];
    double carry = -INFINITY;
    for (long long b = threadIdx.x; b < t; b += kBlock)
        carry = tmax[b] > carry ? tmax[b] : carry;
    carry = block_reduce(carry, OpMaxD(), lds);
    double w[kItems];
    load_tile_blocked(WArray{u}, base, n, w, tile, -INFINITY);
    double m = -INFINITY;
#pragma unroll
    for (int j = 0; j < kItems; ++j)
        m = w[j] > m ? w[j] : m;
    double tot;
    double run = block_exclusive_scan(m, OpMaxD(), -__builtin_huge_val(), lds, &tot);
    run = run > carry ? run : carry;
    const long long ib = base + static_cast<long long>(threadIdx.x) * kItems;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        run = w[j] > run ? w[j] : run;
        if (ib + j < n)
            u[ib + j] = run;
    }
}

// ---------------------------------------------------- residual schemes
//
// The residual transform (transform.hpp:154-168): integer parts
// floor(N W_i), raw residuals modf(N W_i) folded exactly for their total
// (the fast event's summaries and resolver, then the last tile's fold),
// R = N - sum of integer parts draws from the normalised residuals.
// k_fe_resid_tiles gives the raw residuals' tile sums (the fold's
// approximate prefix) and the integer parts' exact tile sums;
// k_fe_resid_params the event's residual parameters, and the normalised
// residuals' approximate tile prefix and predicted crossings for the fast
// event's counting passes.

__global__ void __launch_bounds__(kBlock) k_fe_resid_tiles(WLog wb, long long M,
    double coeff, const ResampleDev *prm, double *tsum, long long *tint)
{
    if (!prm->active)
        return;
    const long long base = static_cast<long long>(blockIdx.x) * kTile;
    if (base >= M)
        return;
    __shared__ double tab[128];
    __shared__ double dl[kWaves];
    __shared__ long long il[kWaves];
    double w[kItems];
    fe_load_w(wb.loaded(), base, M, tab, w);
    double rs = 0.0;
    long long is = 0;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        double ip;
        rs += modf(coeff * w[j], &ip);
        is += static_cast<long long>(ip);
    }
    rs = block_reduce(rs, OpAddD(), dl);
    is = block_reduce(is, OpAddL(), il);
    if (threadIdx.x == 0) {
        tsum[blockIdx.x] = rs;
        tint[blockIdx.x] = is;
    }
}

// One block: the exact total of the M raw residuals (from the last tile's
// exact start: the steps in its binade, else the exact fold) into
// xstart[ntiles]; the sum of the integer parts, R = N - that sum, the total's
// reciprocal (transform.hpp:159-165) and the draw parameters for R draws
// (u01_sequence.hpp: delta = 1 / R, systematic u0 = U delta); then the raw
// residuals' tile prefix rpre scaled to the normalised residuals (into
// tpre) and the tiles predicted to cross a binade.
struct FeResidShared {
    double tile[kTileLds];
    fold::ExactShared ex;
    double tab[128];
    double dlds[kWaves];
    long long il[kWaves];
    double total;
    int cnt;
};

template <class B>
__global__ void __launch_bounds__(kBlock) k_fe_resid_params(WResidRaw<B> src,
    long long M, ResampleDev *prm, double *xstart, const long long *tint,
    const double *rpre, double *tpre, int *fpred, int scheme)
{
    if (!prm->active)
        return;
    const long long nb = (M + kTile - 1) / kTile;
    const long long t = nb - 1;
    const long long base = t * kTile;
    __shared__ FeResidShared sh;
    const int nvalid = static_cast<int>(M - base);
    const double x0 = xstart[t];
    // the integer parts' tile sums load with the last tile
    long long is = 0;
    for (long long b = threadIdx.x; b < nb; b += kBlock)
        is += tint[b];
    double w[kItems];
    fe_load_blocked(src.loaded(), base, M, sh.tile, sh.tab, w);
    const bool xpos = x0 > 0.0 && x0 < INFINITY;
    const FeSteps st(xpos ? fold::binade(x0) : 0);
    double dsum = 0.0;
    bool tie = false;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const double q = st.q(w[j]);
        const double r = FeSteps::round(q);
        tie = tie || FeSteps::half(q, r);
        dsum += r;
    }
    const double D = block_reduce(dsum, OpAddD(), sh.dlds);
    const bool anytie = __syncthreads_or(tie);
    const double a0d = xpos ? static_cast<double>(fold::mant(x0)) : 0.0;
    if (xpos && st.ok && !anytie && a0d + D < static_cast<double>(fold::kTwo53)) {
        if (threadIdx.x == 0)
            sh.total = st.value(a0d + D);
    } else {
        fold::tile_exact(sh.tile, x0, sh.ex);
        if (threadIdx.x == 0)
            sh.total = sh.tile[lds_pad(nvalid - 1)];
    }
    is = block_reduce(is, OpAddL(), sh.il);   // (its barriers publish total)
    const double tot = sh.total;
    if (threadIdx.x == 0) {
        sh.cnt = 0;
        xstart[nb] = tot;
        prm->integ_sum = is;
        const long long nd = prm->nsample - is;
        prm->ndraw = nd;
        prm->resid_total = tot;
        prm->inv_resid = 1 / tot;
        const double delta = 1 / static_cast<double>(nd);
        prm->delta = delta;
        prm->u0 = scheme == kResidualSystematic ?
            u01_u32d(philox_word(prm->key, prm->word0)) * delta : 0.0;
    }
    // the scaled prefix of a thread's consecutive tiles (and the next
    // one's), all in registers, and their predicted crossings
    // (fold::predicted_crossing)
    const double inv = 1 / tot;
    constexpr int kPer = 16;
    const long long b0 = static_cast<long long>(threadIdx.x) * kPer;
    for (long long c0 = 0; c0 < nb; c0 += static_cast<long long>(kBlock) * kPer) {
        double tp[kPer + 1];
#pragma unroll
        for (int k = 0; k <= kPer; ++k) {
            const long long b = c0 + b0 + k;
            tp[k] = rpre[b < nb ? b : nb - 1];
        }
#pragma unroll
        for (int k = 0; k <= kPer; ++k) {
            keep_loaded(tp[k]);
            tp[k] = tp[k] * inv;
        }
        __syncthreads();   // sh.cnt cleared
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const long long b = c0 + b0 + k;
            if (b < nb) {
                tpre[b] = tp[k];
                const bool cross = b == 0 || b + 1 >= nb ||
                    fold::sum_tag(tp[k]) != fold::sum_tag(tp[k + 1]);
                if (cross) {
                    const int q = atomicAdd(&sh.cnt, 1);
                    if (q < kFePred)
                        fpred[1 + q] = static_cast<int>(b);
                }
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0)
        fpred[0] = sh.cnt;
}
