// genealogy.hip -- device ancestry history (ResampleIndex,
// include/vsmc/resample/index.hpp:46-206).
//
// Row t of the history holds the ancestor index of every particle at
// iteration t (the identity for a step that did not resample).
// index(id, t) composes the rows from the last back to t
// (index.hpp:121-134); the index matrix (index.hpp:156-204) does that for
// every particle at once, one thread per particle walking back through the
// rows. An ancestor outside [0, n) stops that particle's walk (the rest of
// its entries are -1) and is reported as an error, instead of being used
// as an address.
#include "genealogy.hpp"

#include <limits.h>

namespace vsmch {

namespace {

__global__ void k_identity_row(int *row, long long n)
{
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        row[i] = static_cast<int>(i);
}

// out(i, t) for t = T-1 .. 0: out(i, T-1) = hist[T-1][i], then
// out(i, t) = hist[t][out(i, t+1)] (index_matrix_dispatch, index.hpp:156-204).
__global__ void k_trace(const int *hist, long long n, long long T, int col_major,
    int *out, int *bad)
{
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int a = static_cast<int>(i);
    for (long long t = T - 1; t >= 0; --t) {
        int v = -1;
        if (a >= 0) {
            v = hist[t * n + a];
            if (v < 0 || v >= n) {
                v = -1;
                atomicOr(bad, 1);
            }
        }
        out[col_major ? t * n + i : i * T + t] = v;
        a = v;
    }
}

// index(id, iter) (index.hpp:121-134): one thread.
__global__ void k_index_one(const int *hist, long long n, long long T,
    long long id, long long iter, int *dscal)
{
    int a = static_cast<int>(id);
    for (long long t = T - 1; t >= iter; --t) {
        const int v = hist[t * n + a];
        if (v < 0 || v >= n) {
            dscal[0] = 1;
            dscal[1] = -1;
            return;
        }
        a = v;
    }
    dscal[0] = 0;
    dscal[1] = a;
}

inline unsigned int blocks_for(long long n)
{
    return static_cast<unsigned int>((n + 255) / 256);
}

// Makes room for one more row (capacity doubles; rows are kept).
int grow(RIndex *r)
{
    if (r->iters < r->cap)
        return kOk;
    const long long cap = r->cap < 4 ? 4 : 2 * r->cap;
    int *h = nullptr;
    if (hipMalloc(&h, sizeof(int) * static_cast<size_t>(cap) * r->n) != hipSuccess) {
        set_error("ResampleIndex: history allocation failed");
        return kErrAlloc;
    }
    if (r->iters > 0)
        VSMCH_HIP(hipMemcpyAsync(h, r->hist,
            sizeof(int) * static_cast<size_t>(r->iters) * r->n,
            hipMemcpyDeviceToDevice, r->st));
    VSMCH_HIP(hipStreamSynchronize(r->st));
    if (r->hist)
        VSMCH_HIP(hipFree(r->hist));
    r->hist = h;
    r->cap = cap;
    return kOk;
}

int write_row(RIndex *r, long long t, const int *idx)
{
    int *row = r->hist + t * r->n;
    if (idx)
        VSMCH_HIP(hipMemcpyAsync(row, idx, sizeof(int) * r->n,
            hipMemcpyDeviceToDevice, r->st));
    else
        hipLaunchKernelGGL(k_identity_row, dim3(blocks_for(r->n)), dim3(256), 0,
            r->st, row, r->n);
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

} // namespace

int rindex_create(RIndex **out, long long n, hipStream_t st)
{
    if (!out || n < 1 || n > INT_MAX) {
        set_error("ResampleIndex: size must be in [1, 2^31 - 1]");
        return kErrInvalid;
    }
    RIndex *r = new RIndex();
    r->n = n;
    r->st = st;
    if (hipMalloc(&r->dscal, sizeof(int) * 2) != hipSuccess) {
        delete r;
        set_error("ResampleIndex: allocation failed");
        return kErrAlloc;
    }
    *out = r;
    return kOk;
}

void rindex_destroy(RIndex *r)
{
    if (!r)
        return;
    if (r->hist)
        (void)hipFree(r->hist);
    if (r->dscal)
        (void)hipFree(r->dscal);
    delete r;
}

// push_back() / push_back(first) (index.hpp:70-87)
int rindex_push_back(RIndex *r, const int *idx)
{
    const int rc = grow(r);
    if (rc != kOk)
        return rc;
    ++r->iters;
    return write_row(r, r->iters - 1, idx);
}

// insert() / insert(first) / insert(iter, first) (index.hpp:89-119);
// iter < 0 stands for the last iteration.
int rindex_insert(RIndex *r, long long iter, const int *idx)
{
    if (iter < 0)
        iter = r->iters - 1;
    if (r->iters == 0 || iter >= r->iters) {
        set_error("ResampleIndex::insert: iteration not recorded");
        return kErrInvalid;
    }
    return write_row(r, iter, idx);
}

// clear() (index.hpp:68): releases the rows; reset() only sets iters = 0.
int rindex_clear(RIndex *r)
{
    if (r->hist)
        VSMCH_HIP(hipFree(r->hist));
    r->hist = nullptr;
    r->cap = 0;
    r->iters = 0;
    return kOk;
}

int rindex_index(RIndex *r, long long id, long long iter, long long *out_host)
{
    if (iter < 0 || iter >= r->iters || id < 0 || id >= r->n) {
        set_error("ResampleIndex::index: particle or iteration out of range");
        return kErrInvalid;
    }
    hipLaunchKernelGGL(k_index_one, dim3(1), dim3(1), 0, r->st, r->hist, r->n,
        r->iters, id, iter, r->dscal);
    VSMCH_HIP(hipGetLastError());
    int h[2];
    VSMCH_HIP(hipMemcpyAsync(h, r->dscal, sizeof(h), hipMemcpyDeviceToHost, r->st));
    VSMCH_HIP(hipStreamSynchronize(r->st));
    if (h[0]) {
        set_error("ResampleIndex::index: an ancestor index is outside [0, N)");
        return kErrInvalid;
    }
    *out_host = h[1];
    return kOk;
}

// index_matrix<RowMajor / ColMajor> (index.hpp:136-148): out holds
// n * iters ints.
int rindex_matrix(RIndex *r, int col_major, int *out)
{
    if (r->iters == 0)
        return kOk;
    VSMCH_HIP(hipMemsetAsync(r->dscal, 0, sizeof(int), r->st));
    hipLaunchKernelGGL(k_trace, dim3(blocks_for(r->n)), dim3(256), 0, r->st,
        r->hist, r->n, r->iters, col_major, out, r->dscal);
    VSMCH_HIP(hipGetLastError());
    int bad = 0;
    VSMCH_HIP(hipMemcpyAsync(&bad, r->dscal, sizeof(int), hipMemcpyDeviceToHost, r->st));
    VSMCH_HIP(hipStreamSynchronize(r->st));
    if (bad) {
        set_error("ResampleIndex::index_matrix: an ancestor index is outside [0, N)");
        return kErrInvalid;
    }
    return kOk;
}

} // namespace vsmch
