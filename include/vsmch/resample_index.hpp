// vsmch/resample_index.hpp -- ResampleIndex (include/vsmc/resample/index.hpp:
// 46-206) for C++ users of the MI355X particle backend: the ancestry
// history lives on the device and is traced there (vsmch_rindex_* in
// include/vsmch.h). Index arrays passed in are device arrays of size()
// ints, e.g. the ancestors vsmch_particle_resample writes to idx_out.
#pragma once

#include "sampler.hpp"

#include <vector>

namespace vsmch {

class ResampleIndex
{
    public:
    using size_type = std::size_t;
    using index_type = int;

    explicit ResampleIndex(size_type N) : size_(N)
    {
        check(vsmch_rindex_create(&r_, static_cast<int64_t>(N)), "ResampleIndex");
    }
    ResampleIndex(const ResampleIndex &) = delete;
    ResampleIndex &operator=(const ResampleIndex &) = delete;
    ~ResampleIndex()
    {
        if (r_)
            vsmch_rindex_destroy(r_);
    }

    size_type size() const { return size_; }

    std::size_t iter_size() const
    {
        int64_t t = 0;
        check(vsmch_rindex_iter_size(r_, &t), "ResampleIndex::iter_size");
        return static_cast<std::size_t>(t);
    }

    void reset() { check(vsmch_rindex_reset(r_), "ResampleIndex::reset"); }
    void clear() { check(vsmch_rindex_clear(r_), "ResampleIndex::clear"); }

    // a step without resampling records the identity
    void push_back() { check(vsmch_rindex_push_back(r_, nullptr), "ResampleIndex::push_back"); }
    void push_back(const index_type *idx_device)
    {
        check(vsmch_rindex_push_back(r_, idx_device), "ResampleIndex::push_back");
    }
    void insert() { check(vsmch_rindex_insert(r_, -1, nullptr), "ResampleIndex::insert"); }
    void insert(const index_type *idx_device)
    {
        check(vsmch_rindex_insert(r_, -1, idx_device), "ResampleIndex::insert");
    }
    void insert(std::size_t iter, const index_type *idx_device)
    {
        check(vsmch_rindex_insert(r_, static_cast<int64_t>(iter), idx_device),
            "ResampleIndex::insert");
    }

    index_type index(size_type id, std::size_t iter) const
    {
        int64_t v = 0;
        check(vsmch_rindex_index(r_, static_cast<int64_t>(id),
                  static_cast<int64_t>(iter), &v),
            "ResampleIndex::index");
        return static_cast<index_type>(v);
    }

    // index_matrix<Layout>() on the host (size() x iter_size() RowMajor,
    // iter_size() x size() ColMajor)
    template <MatrixLayout Layout>
    std::vector<index_type> index_matrix() const
    {
        const std::size_t T = iter_size();
        std::vector<index_type> m(size_ * T);
        if (m.empty())
            return m;
        index_type *d = nullptr;
        check_hip(hipMalloc(&d, sizeof(index_type) * m.size()), "ResampleIndex::index_matrix");
        const vsmch_status s = vsmch_rindex_index_matrix(r_, Layout, d);
        const hipError_t e = hipMemcpy(m.data(), d, sizeof(index_type) * m.size(),
            hipMemcpyDeviceToHost);
        (void)hipFree(d);
        check(s, "ResampleIndex::index_matrix");
        check_hip(e, "ResampleIndex::index_matrix");
        return m;
    }

    // the same into a device array of size() * iter_size() ints
    template <MatrixLayout Layout>
    void read_index_matrix(index_type *out_device) const
    {
        check(vsmch_rindex_index_matrix(r_, Layout, out_device),
            "ResampleIndex::read_index_matrix");
    }

    private:
    size_type size_;
    vsmch_rindex *r_ = nullptr;
};

} // namespace vsmch
