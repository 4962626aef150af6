// genealogy.hpp -- device ancestry history (ResampleIndex,
// include/vsmc/resample/index.hpp:46-206), shared by genealogy.hip and
// capi.hip.
#pragma once

#include "internal.hpp"

namespace vsmch {

// Row t of `hist` (n ints) holds the ancestor of every particle at
// iteration t; rows [0, iters) are in use, `cap` are allocated.
struct RIndex {
    long long n = 0;
    long long iters = 0;
    long long cap = 0;
    int *hist = nullptr;
    int *dscal = nullptr;   // [0]: bad-ancestor flag, [1]: a single index
    hipStream_t st = nullptr;
};

int rindex_create(RIndex **out, long long n, hipStream_t st);
void rindex_destroy(RIndex *r);
int rindex_push_back(RIndex *r, const int *idx);
int rindex_insert(RIndex *r, long long iter, const int *idx);
int rindex_clear(RIndex *r);
int rindex_index(RIndex *r, long long id, long long iter, long long *out_host);
int rindex_matrix(RIndex *r, int col_major, int *out);

} // namespace vsmch
