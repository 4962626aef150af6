// noalloc.hip -- vsmch_resample_trans_u01_rep_noalloc: the replication
// counts of vsmch_resample_trans_u01_rep with every scratch buffer carved
// from one file-scope device array, so that the call allocates no device
// memory (for callers under graph capture or in serving loops that cannot
// call hipMalloc). It runs the same kernels as the allocating entry point.
//
// Limits: m and n at most kNoallocMax (4,194,304). The array, about 33 MiB,
// lives on every device the library runs on until the process ends. All
// calls on one device share it, and it keeps the last call's contents:
// calls that overlap on one device (two host threads, or two streams) race.
#include "internal.hpp"
#include "../../include/vsmch.h"

#include <mutex>
#include <string>
#include <vector>

namespace vsmch {
namespace {

constexpr long long kNoallocMax = 1LL << 22;

constexpr size_t align256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

// The fold's buffer sizes at kNoallocMax weights, as Workspace::ensure
// sizes them.
constexpr long long kNaTiles = (kNoallocMax + kTile - 1) / kTile + 1;
constexpr long long kNaGroups = kNaTiles / kGroup + 2;

// Byte offsets of the buffers trans_u01_rep uses, each 256-byte aligned.
constexpr size_t kOffTsum = 0;
constexpr size_t kOffTpre = kOffTsum + align256(sizeof(double) * kNaTiles);
constexpr size_t kOffSegs = kOffTpre + align256(sizeof(double) * kNaTiles);
constexpr size_t kOffXstart = kOffSegs + align256(sizeof(fold::TileSeg) * kNaTiles);
constexpr size_t kOffXgstart = kOffXstart + align256(sizeof(double) * (kNaTiles + 1));
constexpr size_t kOffGmode = kOffXgstart + align256(sizeof(double) * (kNaGroups + 1));
constexpr size_t kOffGroups = kOffGmode + align256(sizeof(int) * (kNaGroups + 1));
constexpr size_t kOffTagg = kOffGroups + align256(sizeof(GroupSeg) * (kNaGroups + 1));
constexpr size_t kOffPrm = kOffTagg + align256(sizeof(long long) * 3 * kNaTiles);
constexpr size_t kOffV = kOffPrm + align256(sizeof(ResampleDev));
constexpr size_t kScratchBytes = kOffV + align256(sizeof(double) * kNoallocMax);

} // namespace
} // namespace vsmch

// The scratch array. In host code this name is only a stand-in for the
// symbol: the device address comes from hipGetSymbolAddress, per device.
static __device__ __align__(256) unsigned char scratch[vsmch::kScratchBytes];

namespace vsmch {
namespace {

struct NoallocDevice {
    std::once_flag once;
    hipError_t rc = hipSuccess;
    unsigned char *ptr = nullptr;   // this device's copy of `scratch`
};

// Per-device set-up, once per device with that device current.
NoallocDevice *noalloc_device(int dev)
{
    static std::mutex mu;
    static std::vector<NoallocDevice *> devs;   // never freed
    NoallocDevice *d = nullptr;
    {
        std::lock_guard<std::mutex> lk(mu);
        if (devs.size() <= static_cast<size_t>(dev))
            devs.resize(dev + 1, nullptr);
        if (!devs[dev])
            devs[dev] = new NoallocDevice();
        d = devs[dev];
    }
    std::call_once(d->once, [d] {
        void *p = nullptr;
        d->rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        d->ptr = static_cast<unsigned char *>(p);
    });
    return d;
}

// A workspace whose buffers are views into this device's scratch; the
// buffers trans_u01_rep does not use stay null.
Workspace scratch_workspace(unsigned char *base, hipStream_t st)
{
    Workspace ws;
    ws.fixed = true;
    ws.stream = st;
    ws.cap_n = kNoallocMax;
    ws.cap_tiles = kNaTiles;
    ws.tsum = reinterpret_cast<double *>(base + kOffTsum);
    ws.tpre = reinterpret_cast<double *>(base + kOffTpre);
    ws.segs = reinterpret_cast<fold::TileSeg *>(base + kOffSegs);
    ws.xstart = reinterpret_cast<double *>(base + kOffXstart);
    ws.xgstart = reinterpret_cast<double *>(base + kOffXgstart);
    ws.gmode = reinterpret_cast<int *>(base + kOffGmode);
    ws.groups = reinterpret_cast<GroupSeg *>(base + kOffGroups);
    ws.tagg = reinterpret_cast<long long *>(base + kOffTagg);
    ws.prm = reinterpret_cast<ResampleDev *>(base + kOffPrm);
    ws.v = reinterpret_cast<double *>(base + kOffV);
    return ws;
}

} // namespace
} // namespace vsmch

using namespace vsmch;

extern "C" vsmch_status vsmch_resample_trans_u01_rep_noalloc(int m, int n,
    const double *weight, const double *u01, int *replication)
{
    if (m < 0 || n < 0 || (m > 0 && (!weight || !replication)) ||
        (n > 0 && m > 1 && !u01)) {
        set_error("vsmch_resample_trans_u01_rep_noalloc: invalid argument");
        return VSMCH_ERR_INVALID;
    }
    if (m > kNoallocMax || n > kNoallocMax) {
        set_error("vsmch_resample_trans_u01_rep_noalloc: m = " +
            std::to_string(m) + ", n = " + std::to_string(n) +
            " exceed the fixed scratch (at most " +
            std::to_string(kNoallocMax) + " each)");
        return VSMCH_ERR_INVALID;
    }
    int dev = 0;
    VSMCH_HIP(hipGetDevice(&dev));
    NoallocDevice *d = noalloc_device(dev);
    if (d->rc != hipSuccess || d->ptr == nullptr) {
        set_error(std::string("vsmch_resample_trans_u01_rep_noalloc: "
            "hipGetSymbolAddress(scratch): ") + hipGetErrorString(d->rc));
        return VSMCH_ERR_HIP;
    }
    Workspace ws = scratch_workspace(d->ptr, library_stream());
    return trans_u01_rep(ws, m, n, weight, u01, replication);
}
