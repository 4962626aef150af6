// distribution.hpp -- device versions of the vSMC distributions a user's
// eval_sp draws from: U01Distribution (include/vsmc/rng/u01_distribution.hpp:
// 53-73) and NormalDistribution (include/vsmc/rng/normal_distribution.hpp:
// 199-216), with the same word consumption and Box-Muller caching, so a
// device functor draws exactly the numbers the reference functor draws
// (up to libm ulps in log/sqrt/sin/cos).
#pragma once

#include "device_rng.hpp"

namespace vsmch {

template <typename RealType>
class U01Distribution {
  public:
    template <typename RNG>
    __device__ RealType operator()(RNG &rng)
    {
        return u01_from_u32<RealType>(rng.next());
    }
};

template <typename RealType>
class NormalDistribution {
  public:
    __host__ __device__ NormalDistribution(RealType mean = 0,
        RealType stddev = 1)
        : mean_(mean), stddev_(stddev), v_(0), saved_(false)
    {
    }

    template <typename RNG>
    __device__ RealType operator()(RNG &rng)
    {
        RealType z;
        if (saved_) {
            z = v_;
            saved_ = false;
        } else {
            const uint32_t a = rng.next();
            const uint32_t b = rng.next();
            const normal_pair<RealType> p = box_muller<RealType>(a, b);
            z = p.z0;
            v_ = p.z1;
            saved_ = true;
        }
        return rn_add(mean_, rn_mul(stddev_, z));
    }

  private:
    RealType mean_;
    RealType stddev_;
    RealType v_;
    bool saved_;
};

} // namespace vsmch
