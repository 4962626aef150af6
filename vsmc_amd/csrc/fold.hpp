// fold.hpp -- exact emulation of a sequential FP64 left fold on gfx950.
//
// resample_trans_u01_rep (include/vsmc/resample/transform.hpp:61-69) and
// resample_trans_residual (:164) accumulate weights one by one:
// s_i = fl(s_{i-1} + w_i). A parallel scan rounds differently, and at
// N = 2^24 that moves tens of replication counts. This file reproduces the
// sequential values exactly, in parallel:
//
//  * While s stays in one binade [2^e, 2^(e+1)), write s = a * u with
//    u = 2^(e-52) and a an integer in [2^52, 2^53). Adding w >= 0 rounds
//    a + w/u to the nearest integer, ties to even, so one step is the map
//    a -> a + d[a mod 2] with d[0], d[1] depending only on (w, e). These
//    maps compose: (d0, d1) then (c0, c1) is p -> d[p] + c[(p + d[p]) mod 2].
//  * A step whose result leaves the binade (a + d >= 2^53 under the finer
//    rounding, which happens exactly when the real sum rounds to >= 2^(e+1))
//    is done as one real FP64 add from the exact previous value.
//  * A tile (4096 elements) first predicts, from an approximate prefix, where
//    the binade changes; its summary is a short list of (binade, map) runs
//    and crossing weights. One workgroup walks the tile summaries with exact
//    values (64 single-run tiles per wave step), verifying every prediction;
//    a tile whose prediction fails is evaluated element by element.
//  * Each tile then recomputes its own exact prefix from its exact start.
// Weights must be >= 0 and not NaN; +inf is allowed (sums become +inf).
#pragma once

#include "common.hpp"

namespace vsmch {
namespace fold {

constexpr long long kCap = 1LL << 54;
constexpr long long kTwo53 = 1LL << 53;
constexpr int kSegMax = 47;       // crossings a tile summary can hold
constexpr int kEmpty = -100000;   // binade tag of an empty run

struct FMap {
    long long d0;
    long long d1;
};

__host__ __device__ inline FMap ident() { return FMap{0, 0}; }

__host__ __device__ inline long long sat(long long x)
{
    return x >= kCap ? kCap : x;
}

// m1 applied first, then m2.
__host__ __device__ inline FMap compose(FMap m1, FMap m2)
{
    FMap r;
    r.d0 = m1.d0 >= kCap ? kCap : sat(m1.d0 + ((m1.d0 & 1) ? m2.d1 : m2.d0));
    r.d1 = m1.d1 >= kCap ? kCap :
                           sat(m1.d1 + (((m1.d1 + 1) & 1) ? m2.d1 : m2.d0));
    return r;
}

struct OpCompose {
    __device__ FMap operator()(FMap a, FMap b) const { return compose(a, b); }
};

// Saturating sum of map offsets (both <= kCap, so no overflow).
struct OpAddSatL {
    __device__ long long operator()(long long a, long long b) const
    {
        return sat(a + b);
    }
};

// Binade exponent of x > 0 (finite); subnormals share the lowest normal ulp.
__host__ __device__ inline int binade(double x)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    int eb = static_cast<int>((b >> 52) & 0x7FF);
    return (eb == 0 ? 1 : eb) - 1023;
}

// Integer a with x = a * 2^(binade(x) - 52).
__host__ __device__ inline long long mant(double x)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const long long frac = static_cast<long long>(b & ((1ULL << 52) - 1));
    return ((b >> 52) & 0x7FF) == 0 ? frac : (frac | (1LL << 52));
}

__host__ __device__ inline double from_mant(long long a, int e)
{
    return ldexp(static_cast<double>(a), e - 52);
}

// Map of adding w (>= 0) while the sum stays in binade e.
__host__ __device__ inline FMap elem_map(double w, int e)
{
    const double q = ldexp(w, 52 - e);
    if (!(q < static_cast<double>(kCap)))
        return FMap{kCap, kCap};
    const long long k = static_cast<long long>(q);
    const double f = q - static_cast<double>(k);
    if (f < 0.5)
        return FMap{k, k};
    if (f > 0.5)
        return FMap{k + 1, k + 1};
    // tie: pick the even result for each parity of a
    return FMap{(k & 1) ? k + 1 : k, (k & 1) ? k : k + 1};
}

struct TileSeg {
    int ncross;              // crossings in the tile; -1: evaluate exactly
    int e[kSegMax + 1];      // binade of each run (kEmpty if empty)
    FMap map[kSegMax + 1];   // composed map of each run
    double cross[kSegMax];   // weight of each crossing element
};

// Applies a tile summary to the exact start x. Returns false when a
// prediction does not hold (the caller then evaluates the tile exactly).
__device__ inline bool apply_segments(const TileSeg &sg, double &x)
{
    if (sg.ncross < 0)
        return false;
    for (int r = 0; r <= sg.ncross; ++r) {
        if (sg.e[r] != kEmpty) {
            if (!(x > 0.0) || !(x < INFINITY))
                return false;
            if (binade(x) != sg.e[r])
                return false;
            const long long a0 = mant(x);
            const long long a = a0 + ((a0 & 1) ? sg.map[r].d1 : sg.map[r].d0);
            if (a >= kTwo53)
                return false;
            x = from_mant(a, sg.e[r]);
        }
        if (r < sg.ncross)
            x = x + sg.cross[r];
    }
    return true;
}

// Runs of one tile's multi-run pass (run starts at possible crossings).
constexpr int kMaxRuns = 64;
constexpr int kZeroTag = -4000;    // sum_tag of 0
constexpr int kInfTag = 4000;      // sum_tag of +inf
constexpr int kNoTag = -8000;      // run of one element
// Relative distance between an FP64 prefix of a tile (exact start + block
// scan) and the sequential fold: at most (kTile + 32) * 2^-53 < 2^-41;
// twice that.
constexpr double kSumMargin = 0x1p-40;

// Segmented (run count, rounded-step sum) pairs: a run start resets the sum.
struct SegCnt {
    int cnt;
    double d;
};
struct OpSegCnt {
    __device__ SegCnt operator()(SegCnt a, SegCnt b) const
    {
        return SegCnt{a.cnt + b.cnt, b.cnt ? b.d : a.d + b.d};
    }
};

// Segmented composition of (flag, map) pairs: a set flag starts a run.
struct SegPair {
    long long flag;
    FMap m;
};
struct OpSeg {
    __device__ SegPair operator()(SegPair a, SegPair b) const
    {
        SegPair r;
        r.flag = a.flag | b.flag;
        r.m = b.flag ? b.m : compose(a.m, b.m);
        return r;
    }
};

constexpr int kMaxTies = 32;   // half-ulp ties a one-run tile can carry

struct SingleShared {
    double dred[kWaves];
    int nred[kWaves];
    int tk[kMaxTies];       // element of each tie, in order
    double tpre[kMaxTies];  // rounded steps before it
    double tq[kMaxTies];    // its step without the half
    int cum[kMaxTies];      // extra ulps from the ties up to it
    int ok;
};

struct OpAddI {
    __device__ int operator()(int a, int b) const { return a + b; }
};

struct ExactShared {
    SingleShared single;
    FMap scan[kWaves];
    int red[kWaves];
    double dred[kWaves];
    SegCnt sred[kWaves];
    SegPair pred[kWaves];
    int first_start[kBlock];       // thread's first element starts a run
    int rf[kMaxRuns + 1];          // first element of each run
    int rpe[kMaxRuns];             // binade tag of the run after its start
    FMap rM[kMaxRuns];             // map of the run after its start
    double rs[kMaxRuns];           // exact sum at the run's start
    int pos;                       // first element not yet folded
    double cur;                    // exact sum before it
};

// Binade of a sum, with tags of their own for 0 and +inf.
__device__ inline int sum_tag(double s)
{
    if (s == 0.0)
        return kZeroTag;
    if (!(s < INFINITY))
        return kInfTag;
    return binade(s);
}

// Element k may start a binade: the smallest value the exact s_{k-1} can
// take and the largest value s_k can take lie in different binades.
__device__ inline bool may_cross(double sprev, double s)
{
    return sum_tag(sprev * (1.0 - kSumMargin)) != sum_tag(s * (1.0 + kSumMargin));
}

__device__ inline bool finite_tag(int tg) { return tg != kZeroTag && tg != kInfTag; }

// Element idx starts a run of the multi-run pass: the pass's first element,
// or an add that may change binade. An add of 0 leaves the sum exactly as
// it is, so it never starts one (the zero padding after a partial tile's
// last weight would otherwise make a run of every padding element when the
// total sits at a binade boundary, as the normalised weights' total 1 does).
__device__ inline bool run_start(int idx, int pos0, double w, double sprev,
    double s)
{
    return idx == pos0 || (w > 0.0 && may_cross(sprev, s));
}

// Rounded-step total as a map (saturating: at or above 2^54 the run
// certainly leaves its binade).
__device__ inline FMap steps_map(double d)
{
    const long long k = d < static_cast<double>(kCap) ? static_cast<long long>(d) : kCap;
    return FMap{k, k};
}

// Multi-run pass of tile_exact over elements pos0.. (the ones before hold
// their sums), from the exact sum cur0 before element pos0. An FP64 prefix
// marks every element where the sum may change binade as a run start
// (s = s_prev + w there: one real add). Inside a run of binade e the fold is
// a composition of the per-element maps of adding w at ulp 2^(e-52) (the
// prefix sum of rint(w / ulp) when no half-ulp tie occurs), so one
// segmented scan gives every run's map and a serial chain over the runs
// gives each run's exact start, verifying that it lies in the predicted
// binade and that the run ends below 2^53. The sums of the verified runs
// are written in place; sh.pos / sh.cur receive the first element left and
// the exact sum before it (kTile when the tile is done).
__device__ __forceinline__ void tile_runs(double *t, int pos0, double cur0,
    ExactShared &sh)
{
    const int base = threadIdx.x * kItems;
    double *const mine = t + lds_pad(base);
    double wsum = 0.0;
#pragma unroll 1
    for (int j = 0; j < kItems; ++j)
        if (base + j >= pos0)
            wsum += mine[j];
    double wtot;
    const double wpre = block_exclusive_scan(wsum, OpAddD(), 0.0, sh.dred, &wtot);
    // pass 1: run starts and rounded steps; ties
    SegCnt loc{0, 0.0};
    bool tie = false;
    {
        double acc = wpre;
        double sprev = cur0 + wpre;
#pragma unroll 1
        for (int j = 0; j < kItems; ++j) {
            const int idx = base + j;
            bool start = false;
            if (idx >= pos0) {
                const double w = mine[j];
                acc += w;
                const double s = cur0 + acc;
                start = run_start(idx, pos0, w, sprev, s);
                if (start) {
                    loc.cnt += 1;
                    loc.d = 0.0;
                } else {
                    const int tg = sum_tag(s);
                    if (finite_tag(tg)) {
                        const double q = ldexp(w, 52 - tg);
                        tie = tie || (q - trunc(q) == 0.5);
                        loc.d += rint(q);
                    }
                }
                sprev = s;
            }
            if (j == 0)
                sh.first_start[threadIdx.x] = start ? 1 : 0;
        }
    }
    SegCnt tot;
    const SegCnt carry = block_exclusive_scan(loc, OpSegCnt(), SegCnt{0, 0.0},
        sh.sred, &tot);
    const bool anytie = __syncthreads_or(tie);
    // with a tie the steps depend on the parity of the sum: compose maps
    SegPair mcarry{0, ident()};
    if (anytie) {
        SegPair lm{0, ident()};
        double acc = wpre;
        double sprev = cur0 + wpre;
#pragma unroll 1
        for (int j = 0; j < kItems; ++j) {
            const int idx = base + j;
            if (idx >= pos0) {
                const double w = mine[j];
                acc += w;
                const double s = cur0 + acc;
                if (run_start(idx, pos0, w, sprev, s)) {
                    lm.flag = 1;
                    lm.m = ident();
                } else {
                    const int tg = sum_tag(s);
                    if (finite_tag(tg))
                        lm.m = compose(lm.m, elem_map(w, tg));
                }
                sprev = s;
            }
        }
        SegPair mtot;
        mcarry = block_exclusive_scan(lm, OpSeg(), SegPair{0, ident()}, sh.pred,
            &mtot);
    }
    // pass 2: each run's first element, and at its last element the binade
    // tag and map
    {
        int r = carry.cnt - 1;
        double dacc = carry.d;
        FMap macc = mcarry.m;
        double acc = wpre;
        double sprev = cur0 + wpre;
        bool pstart = false;
        int ptg = kNoTag;
#pragma unroll 1
        for (int j = 0; j < kItems; ++j) {
            const int idx = base + j;
            if (idx >= pos0) {
                const double w = mine[j];
                acc += w;
                const double s = cur0 + acc;
                const bool start = run_start(idx, pos0, w, sprev, s);
                int tg = kNoTag;
                if (start) {
                    if (idx > base && idx > pos0 && r >= 0 && r < kMaxRuns) {
                        sh.rM[r] = anytie ? macc : steps_map(dacc);
                        sh.rpe[r] = pstart ? kNoTag : ptg;
                    }
                    ++r;
                    dacc = 0.0;
                    macc = ident();
                    if (r <= kMaxRuns)
                        sh.rf[r] = idx;
                } else {
                    tg = sum_tag(s);
                    if (finite_tag(tg)) {
                        if (anytie)
                            macc = compose(macc, elem_map(w, tg));
                        else
                            dacc += rint(ldexp(w, 52 - tg));
                    }
                }
                pstart = start;
                ptg = tg;
                sprev = s;
            }
        }
        const bool ends = base + kItems - 1 >= pos0 &&
            (threadIdx.x == kBlock - 1 || sh.first_start[threadIdx.x + 1] != 0);
        if (ends && r >= 0 && r < kMaxRuns) {
            sh.rM[r] = anytie ? macc : steps_map(dacc);
            sh.rpe[r] = pstart ? kNoTag : ptg;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nr = tot.cnt < kMaxRuns ? tot.cnt : kMaxRuns;
        double x = cur0;
        int ok = 0;
        for (; ok < nr; ++ok) {
            const double s = x + t[lds_pad(sh.rf[ok])];
            const int pe = sh.rpe[ok];
            if (pe != kNoTag) {
                if (sum_tag(s) != pe)
                    break;
                if (finite_tag(pe)) {
                    const long long a0 = mant(s);
                    const FMap m = sh.rM[ok];
                    const long long a = a0 + ((a0 & 1) ? m.d1 : m.d0);
                    if (a >= kTwo53)
                        break;
                    sh.rs[ok] = s;
                    x = from_mant(a, pe);
                    continue;
                }
            }
            sh.rs[ok] = s;
            x = s;
        }
        sh.pos = ok < tot.cnt ? sh.rf[ok] : kTile;
        sh.cur = x;
        sh.red[0] = ok;
    }
    __syncthreads();
    const int nok = sh.red[0];
    // pass 3: the sums of the verified runs
    {
        int r = carry.cnt - 1;
        double dacc = carry.d;
        FMap macc = mcarry.m;
        double acc = wpre;
        double sprev = cur0 + wpre;
#pragma unroll 1
        for (int j = 0; j < kItems; ++j) {
            const int idx = base + j;
            if (idx >= pos0) {
                const double w = mine[j];
                acc += w;
                const double s = cur0 + acc;
                const bool start = run_start(idx, pos0, w, sprev, s);
                sprev = s;
                if (start) {
                    ++r;
                    dacc = 0.0;
                    macc = ident();
                    if (r < nok)
                        mine[j] = sh.rs[r];
                } else if (r < nok) {
                    const int tg = sum_tag(s);
                    const double x = sh.rs[r];
                    if (finite_tag(tg)) {
                        const long long a0 = mant(x);
                        double a;
                        if (anytie) {
                            macc = compose(macc, elem_map(w, tg));
                            a = static_cast<double>(a0 + ((a0 & 1) ? macc.d1 : macc.d0));
                        } else {
                            dacc += rint(ldexp(w, 52 - tg));
                            a = static_cast<double>(a0) + dacc;
                        }
                        mine[j] = ldexp(a, tg - 52);
                    } else {
                        mine[j] = x;
                    }
                }
            }
        }
    }
    __syncthreads();
}

// The common tile: every sum stays in the binade e of the exact start x0.
// Then s_k = x0 advanced by the prefix sum of the rounded steps
// rint(w / 2^(e-52)), and whether that holds is known from the total (the
// sums only grow). A step of exactly f + 1/2 ulp (a tie) rounds to the even
// one of a + f and a + f + 1, which depends on the running sum; up to
// kMaxTies of them are resolved in order by one thread. Writes the tile
// and returns true, or writes nothing and returns false. All threads call
// it.
__device__ __forceinline__ bool tile_single(double *t, double x0, SingleShared &sh)
{
    const int e = sum_tag(x0);
    if (!finite_tag(e))
        return false;
    const int base = threadIdx.x * kItems;
    double *const mine = t + lds_pad(base);
    double dsum = 0.0;
    int ntie = 0;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const double q = ldexp(mine[j], 52 - e);
        const double f = trunc(q);
        const bool tie = q - f == 0.5;
        dsum += tie ? f : rint(q);
        ntie += tie ? 1 : 0;
    }
    double dtot;
    double run = block_exclusive_scan(dsum, OpAddD(), 0.0, sh.dred, &dtot);
    const double a0d = static_cast<double>(mant(x0));
    const double two53 = static_cast<double>(kTwo53);
    if (!__syncthreads_or(ntie > 0)) {
        if (!(a0d + dtot < two53))
            return false;
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            run += rint(ldexp(mine[j], 52 - e));
            mine[j] = ldexp(a0d + run, e - 52);
        }
        __syncthreads();
        return true;
    }
    int ntot;
    const int toff = block_exclusive_scan(ntie, OpAddI(), 0, sh.nred, &ntot);
    if (ntot > kMaxTies || !(a0d + dtot < two53))
        return false;
    if (ntie > 0) {
        double r = run;
        int q_i = toff;
#pragma unroll 1
        for (int j = 0; j < kItems; ++j) {
            const double q = ldexp(mine[j], 52 - e);
            const double f = trunc(q);
            if (q - f == 0.5) {
                sh.tk[q_i] = base + j;
                sh.tpre[q_i] = r;
                sh.tq[q_i] = f;
                ++q_i;
                r += f;
            } else {
                r += rint(q);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // every step total is below 2^53 - a0 here, so exact as an integer
        const long long a0 = mant(x0);
        long long adj = 0;
        for (int i = 0; i < ntot; ++i) {
            const long long ab = a0 + static_cast<long long>(sh.tpre[i]) + adj;
            adj += ((ab + static_cast<long long>(sh.tq[i])) & 1) ? 1 : 0;
            sh.cum[i] = static_cast<int>(adj);
        }
        sh.ok = a0 + static_cast<long long>(dtot) + adj < kTwo53 ? 1 : 0;
    }
    __syncthreads();
    if (!sh.ok)
        return false;
    int ti = toff;
    double adj = ti > 0 ? static_cast<double>(sh.cum[ti - 1]) : 0.0;
#pragma unroll 1
    for (int j = 0; j < kItems; ++j) {
        const double q = ldexp(mine[j], 52 - e);
        const double f = trunc(q);
        if (q - f == 0.5) {
            run += f;
            adj = static_cast<double>(sh.cum[ti]);
            ++ti;
        } else {
            run += rint(q);
        }
        mine[j] = ldexp(a0d + run + adj, e - 52);
    }
    __syncthreads();
    return true;
}

// Element-by-element exact fold of one tile from element pos, whose
// preceding exact sum is cur (the elements before pos hold their sums).
// All threads of the block must call it (it synchronises, also before
// returning).
__device__ __forceinline__ void tile_exact_from(double *t, int pos, double cur,
    ExactShared &sh)
{
    const int base = threadIdx.x * kItems;
    double *const mine = t + lds_pad(base);
    while (pos < kTile) {
        if (!(cur < INFINITY)) {
#pragma unroll 1
            for (int j = 0; j < kItems; ++j)
                if (base + j >= pos)
                    mine[j] = cur;
            break;
        }
        if (cur == 0.0) {
            // the sum stays 0 up to the first non-zero weight k; s_k = w_k
            int lf = kTile;
#pragma unroll 1
            for (int j = 0; j < kItems; ++j)
                if (base + j >= pos && lf == kTile && mine[j] != 0.0)
                    lf = base + j;
            const int k = block_reduce(lf, OpMinI(), sh.red);
#pragma unroll 1
            for (int j = 0; j < kItems; ++j) {
                const int idx = base + j;
                if (idx >= pos && idx < k)
                    mine[j] = 0.0;
            }
            if (k >= kTile)
                break;
            cur = t[lds_pad(k)];
            pos = k + 1;
            continue;
        }
        const int e = binade(cur);
        const long long a0 = mant(cur);
        const int p0 = static_cast<int>(a0 & 1);
        // Without an exact half-ulp tie the maps do not depend on parity:
        // the fold is a prefix sum of the rounded steps d_j = rint(w_j /
        // ulp). They are integers, exact in FP64 below 2^53, and a sum at or
        // above 2^53 (exact or not) marks the binade crossing, so this path
        // runs in FP64 arithmetic (a non-finite sum also counts as a
        // crossing, which the real add then carries).
        const double a0d = static_cast<double>(a0);
        const double two53 = static_cast<double>(kTwo53);
        double dsum = 0.0;
        bool tie = false;
#pragma unroll 1
        for (int j = 0; j < kItems; ++j) {
            if (base + j >= pos) {
                const double q = ldexp(mine[j], 52 - e);
                tie = tie || (q - trunc(q) == 0.5);
                dsum += rint(q);
            }
        }
        int lc = kTile;
        if (!__syncthreads_or(tie)) {
            double dtot;
            double run = block_exclusive_scan(dsum, OpAddD(), 0.0, sh.dred, &dtot);
#pragma unroll 1
            for (int j = 0; j < kItems; ++j) {
                const int idx = base + j;
                if (idx >= pos) {
                    run += rint(ldexp(mine[j], 52 - e));
                    const double a = a0d + run;
                    if (!(a < two53)) {
                        if (lc == kTile)
                            lc = idx;
                    } else {
                        mine[j] = ldexp(a, e - 52);
                    }
                }
            }
        } else {
            FMap magg = ident();
#pragma unroll 1
            for (int j = 0; j < kItems; ++j)
                if (base + j >= pos)
                    magg = compose(magg, elem_map(mine[j], e));
            FMap total;
            FMap run = block_exclusive_scan(magg, OpCompose(), ident(), sh.scan,
                &total);
#pragma unroll 1
            for (int j = 0; j < kItems; ++j) {
                const int idx = base + j;
                if (idx >= pos) {
                    run = compose(run, elem_map(mine[j], e));
                    const long long a = a0 + (p0 ? run.d1 : run.d0);
                    if (a >= kTwo53) {
                        if (lc == kTile)
                            lc = idx;
                    } else {
                        mine[j] = from_mant(a, e);
                    }
                }
            }
        }
        // the sums are non-decreasing, so every element from the first
        // crossing c on is still a weight
        const int c = block_reduce(lc, OpMinI(), sh.red);
        if (c >= kTile)
            break;
        // the crossing element: one real add from the exact previous sum
        const double prev = (c == pos) ? cur : t[lds_pad(c - 1)];
        cur = prev + t[lds_pad(c)];
        __syncthreads();
        if (c >= base && c < base + kItems)
            mine[c - base] = cur;
        pos = c + 1;
    }
    __syncthreads();
}

// Exact inclusive fold of one tile, in place: t (element k at lds_pad(k))
// holds the weights on entry and s_k on return, from the exact start x0.
// Multi-run passes fold the tile (another pass continues after
// kMaxRuns runs or a failed prediction); the element-by-element fold
// finishes from a run whose prediction fails twice. All threads
// of the block must call it (it synchronises, also before returning).
__device__ __forceinline__ void tile_exact(double *t, double x0, ExactShared &sh)
{
    if (tile_single(t, x0, sh.single))
        return;
    int pos = 0;
    double cur = x0;
    while (pos < kTile) {
        const int p0 = pos;
        tile_runs(t, p0, cur, sh);
        pos = sh.pos;
        cur = sh.cur;
        if (pos == p0) {
            tile_exact_from(t, pos, cur, sh);
            break;
        }
    }
}

// Whether tile t (of ntiles) of an approximate tile prefix tpre is
// predicted to need more than a one-binade summary: tile 0 (sums from 0),
// the last tile (its end, the total, lies at a binade boundary), or
// approximate start and end in different binades.
__device__ inline bool predicted_crossing(const double *tpre, long long t,
    long long ntiles)
{
    if (t == 0 || t + 1 >= ntiles)
        return true;
    return sum_tag(tpre[t]) != sum_tag(tpre[t + 1]);
}

struct SegShared {
    double tile[kTileLds];
    double dscan[kWaves];
    long long lscan[kWaves];
    FMap fscan[kWaves];
    SegPair pscan[kWaves];
    double tlast[kBlock];
    unsigned short cross[kBlock];   // crossing flags, bit j of entry t: element t*kItems+j
    TileSeg seg;
};

__device__ inline int prev_tag(double prev)
{
    return (prev > 0.0 && prev < INFINITY) ? binade(prev) : kEmpty;
}

// Tile summary with crossings (the rare case): the approximate running sums
// from tpre + ex (ex: this thread's exclusive share of the block scan) mark
// the elements whose add changes binade; runs between them carry the
// composition of their elements' maps. Per-element values are recomputed in
// each pass from the same running sums.
__device__ __forceinline__ void tile_segments_multi(int nvalid, double tpre,
    double ex, TileSeg *out, SegShared &sh)
{
    const int base = threadIdx.x * kItems;
    const double *const mine = sh.tile + lds_pad(base);
    const double start = tpre + ex;
    unsigned mask = 0;
    double prev = start;
#pragma unroll 1
    for (int j = 0; j < kItems; ++j) {
        const double t = prev + mine[j];
        if (base + j < nvalid) {
            const bool c = !(prev > 0.0) || !(prev < INFINITY) || !(t < INFINITY) ||
                binade(t) != binade(prev);
            mask |= c ? (1u << j) : 0u;
        }
        prev = t;
    }
    // The block scan and the previous thread's running sum approximate the
    // same prefix with different roundings; where they disagree on the
    // binade, this thread's first element is taken as a crossing so that
    // no run mixes binades (each run's binade is verified at its start).
    sh.tlast[threadIdx.x] = prev;
    __syncthreads();
    if (threadIdx.x > 0 && base < nvalid && !(mask & 1u) &&
        prev_tag(sh.tlast[threadIdx.x - 1]) != prev_tag(start))
        mask |= 1u;
    sh.cross[threadIdx.x] = static_cast<unsigned short>(mask);
    long long ltot;
    long long rank = block_exclusive_scan(static_cast<long long>(__popc(mask)),
        OpAddL(), 0LL, sh.lscan, &ltot);
    if (threadIdx.x == 0) {
        sh.seg.ncross = ltot > kSegMax ? -1 : static_cast<int>(ltot);
        for (int r = 0; r <= kSegMax; ++r) {
            sh.seg.e[r] = kEmpty;
            sh.seg.map[r] = ident();
        }
        for (int r = 0; r < kSegMax; ++r)
            sh.seg.cross[r] = 0.0;
    }
    __syncthreads();
    auto cross_at = [&](int k) {
        return k >= nvalid || ((sh.cross[k / kItems] >> (k % kItems)) & 1u) != 0;
    };
    // segmented scan: a run starts at element 0, at a crossing, and right
    // after a crossing; crossing elements carry the identity
    auto pair_at = [&](int j, double pv) {
        const int idx = base + j;
        const bool cr = ((mask >> j) & 1u) != 0;
        const bool prevc = idx == 0 ? true : cross_at(idx - 1);
        const int eb = prev_tag(pv);
        SegPair p;
        p.flag = (prevc || cr) ? 1 : 0;
        p.m = (cr || idx >= nvalid || eb == kEmpty) ? ident() : elem_map(mine[j], eb);
        return p;
    };
    SegPair agg{0, ident()};
    prev = start;
#pragma unroll 1
    for (int j = 0; j < kItems; ++j) {
        agg = OpSeg()(agg, pair_at(j, prev));
        prev += mine[j];
    }
    SegPair stot;
    SegPair run = block_exclusive_scan(agg, OpSeg(), SegPair{0, ident()},
        sh.pscan, &stot);
    if (ltot <= kSegMax) {
        prev = start;
#pragma unroll 1
        for (int j = 0; j < kItems; ++j) {
            const int idx = base + j;
            run = OpSeg()(run, pair_at(j, prev));
            if (idx < nvalid) {
                if ((mask >> j) & 1u) {
                    sh.seg.cross[rank] = mine[j];
                    ++rank;
                } else if (cross_at(idx + 1)) {
                    sh.seg.e[rank] = prev_tag(prev);
                    sh.seg.map[rank] = run.m;
                }
            }
            prev += mine[j];
        }
    }
    __syncthreads();
    const int words = static_cast<int>(sizeof(TileSeg) / 4);
    const int *src = reinterpret_cast<const int *>(&sh.seg);
    int *dst = reinterpret_cast<int *>(out);
    for (int i = threadIdx.x; i < words; i += kBlock)
        dst[i] = src[i];
    __syncthreads();
}

// Tile summary from the approximate prefix tpre (start of this tile); the
// tile is in sh.tile (element k at lds_pad(k)), nvalid elements valid.
// Common case: the approximate start and end of the tile lie in one binade,
// so the summary is one run whose map is the composition of the elements'
// maps (the prefix sum of rounded steps when no half-ulp tie occurs). Only
// ncross, e[0] and map[0] are written then; the resolver verifies every
// prediction against the exact sums.
__device__ __forceinline__ void tile_segments(int nvalid, double tpre,
    TileSeg *out, SegShared &sh)
{
    const int base = threadIdx.x * kItems;
    const double *const mine = sh.tile + lds_pad(base);
    double tsum = 0.0;
#pragma unroll
    for (int j = 0; j < kItems; ++j)
        tsum += mine[j];
    double dtot;
    const double ex = block_exclusive_scan(tsum, OpAddD(), 0.0, sh.dscan, &dtot);
    const double end = tpre + dtot;
    if (!(tpre > 0.0) || !(end < INFINITY) || binade(tpre) != binade(end)) {
        tile_segments_multi(nvalid, tpre, ex, out, sh);
        return;
    }
    const int e = binade(tpre);
    double dsum = 0.0;
    bool tie = false;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        if (base + j < nvalid) {
            const double q = ldexp(mine[j], 52 - e);
            tie = tie || (q - trunc(q) == 0.5);
            dsum += rint(q);
        }
    }
    FMap m;
    if (!__syncthreads_or(tie)) {
        // rounded steps are integers, exact in FP64 below 2^53 (larger or
        // non-finite sums saturate: the resolver then sees the crossing)
        double dd;
        block_exclusive_scan(dsum, OpAddD(), 0.0, sh.dscan, &dd);
        const long long d = dd < static_cast<double>(kCap) ? static_cast<long long>(dd) : kCap;
        m = FMap{d, d};
    } else {
        FMap agg = ident();
#pragma unroll 1
        for (int j = 0; j < kItems; ++j)
            if (base + j < nvalid)
                agg = compose(agg, elem_map(mine[j], e));
        block_exclusive_scan(agg, OpCompose(), ident(), sh.fscan, &m);
    }
    if (threadIdx.x == 0) {
        out->ncross = 0;
        out->e[0] = e;
        out->map[0] = m;
    }
}

} // namespace fold
} // namespace vsmch
