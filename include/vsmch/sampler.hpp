// sampler.hpp -- vSMC's Sampler / Particle / Weight / Monitor API over the
// MI355X particle backend (libvsmch, include/vsmch.h).
//
// A user TU compiled with hipcc keeps the reference's programming model:
//
//   vsmch::Seed::instance().set(101);
//   vsmch::Sampler<MyState> sampler(N, vsmch::Stratified, 0.5);
//   sampler.init(MyInit()).move(MyMove(), false).monitor("pos", 2, MyEval());
//   sampler.initialize(param).iterate(T - 1);
//   std::cout << sampler;
//
// where MyState derives from StateHIP<Dim, Real> and the functors derive
// from InitializeHIP / MoveHIP / MonitorEvalHIP (the CRTP shapes of
// include/vsmc/smp/backend_base.hpp:92-494 and backend_omp.hpp:64-133) with
// a __device__ eval_sp. eval_param / eval_pre / eval_post stay on the host.
// Weights, resampling (counts bit-exact with the reference), the in-place
// gather and the Monitor reduction run in the library.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <iomanip>
#include <iterator>
#include <limits>
#include <map>
#include <memory>
#include <ostream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../vsmch.h"
#include "device_rng.hpp"
#include "distribution.hpp"

namespace vsmch {

// Errors of the C ABI surface as exceptions (the reference throws
// vsmc::RuntimeAssert with VSMC_RUNTIME_ASSERT_AS_EXCEPTION,
// include/vsmc/internal/assert.hpp:42-61).
class RuntimeError : public std::runtime_error {
  public:
    explicit RuntimeError(const std::string &msg) : std::runtime_error(msg) {}
};

inline void check(vsmch_status s, const char *what)
{
    if (s != VSMCH_OK)
        throw RuntimeError(std::string(what) + ": " + vsmch_last_error());
}

inline void check_hip(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        throw RuntimeError(std::string(what) + ": " + hipGetErrorString(e));
}

// include/vsmc/internal/defines.hpp:49-56
enum ResampleScheme {
    Multinomial,
    Stratified,
    Systematic,
    Residual,
    ResidualStratified,
    ResidualSystematic
};

// include/vsmc/core/monitor.hpp:54-58
enum MonitorStage { MonitorMove, MonitorResample, MonitorMCMC };

// SeedGenerator<NullType, unsigned> (include/vsmc/rng/seed.hpp:84-228):
// get() = skip() then the seed; a Particle of size N takes N + 1 seeds.
class Seed {
  public:
    static Seed &instance()
    {
        static Seed s;
        return s;
    }
    void set(uint32_t s) { seed_ = s % max_; }
    uint32_t seed() const { return seed_; }
    uint32_t get()
    {
        seed_ = seed_ < max_ ? seed_ + 1 : 1;
        return seed_;
    }
    // Reserves n consecutive seeds; returns the state before the first.
    uint32_t reserve(uint64_t n)
    {
        const uint32_t base = seed_;
        seed_ = static_cast<uint32_t>((static_cast<uint64_t>(seed_) + n) % max_);
        if (seed_ == 0 && n > 0)
            seed_ = max_;
        return base;
    }
    Seed(const Seed &) = delete;
    Seed &operator=(const Seed &) = delete;

  private:
    Seed() = default;
    uint32_t seed_ = 0;
    uint32_t max_ = 0xFFFFFFFFu;
};

template <typename T>
class Particle;

// MatrixLayout (include/vsmc/internal/defines.hpp), with the C API's values.
enum MatrixLayout { RowMajor = VSMCH_ROW_MAJOR, ColMajor = VSMCH_COL_MAJOR };

// ColMajor StateMatrix<Dim, RealType> in device memory
// (include/vsmc/core/state_matrix.hpp:319-428): component pos of particle id
// at data[pos * N + id].
template <std::size_t Dim, typename RealType>
class StateHIP {
  public:
    using size_type = std::size_t;
    using state_type = RealType;

    struct view_type {
        RealType *data;
        std::size_t n;
        __device__ RealType &operator()(std::size_t id, std::size_t pos) const
        {
            return data[pos * n + id];
        }
    };

    explicit StateHIP(size_type N) : n_(N) {}
    static constexpr std::size_t dim() { return Dim; }

    // StateMatrix::copy (state_matrix.hpp:246-253): particle dst takes the
    // state particle index[dst] had before the call; index is a host array
    // (any integer type, as Particle::resample passes it).
    template <typename IntType>
    void copy(size_type N, const IntType *index)
    {
        if (index == nullptr) {
            resample_event(N);
            return;
        }
        if (N != n_)
            throw std::runtime_error("StateHIP::copy: N differs from the size");
        std::vector<int> h(N);
        for (size_type i = 0; i != N; ++i)
            h[i] = static_cast<int>(index[i]);
        int *d = nullptr;
        check_hip(hipMalloc(&d, sizeof(int) * (N ? N : 1)), "StateHIP::copy");
        check_hip(hipMemcpy(d, h.data(), sizeof(int) * N, hipMemcpyHostToDevice),
            "StateHIP::copy");
        const vsmch_status rc = vsmch_particle_copy(h_, static_cast<int64_t>(N), d);
        (void)hipFree(d);
        check(rc, "StateHIP::copy");
    }
    // copy(N, nullptr): what Particle::resample calls when the weight class
    // has no host weights (resample_data() == nullptr, particle.hpp:175-192)
    // once it has decided to resample: the whole event on the device --
    // counts with the state's scheme from the particle set's resampling
    // engine, ancestors, in-place gather and equal weights.
    void copy(size_type N, std::nullptr_t) { resample_event(N); }

    // The scheme copy(N, nullptr) resamples with (Sampler sets it; the
    // reference Sampler's default is Stratified, sampler.hpp:83-114).
    void resample_scheme(int scheme) { scheme_ = scheme; }
    int resample_scheme() const { return scheme_; }
    size_type size() const { return n_; }
    view_type view() const { return view_type{data_, n_}; }
    RealType *data() const { return data_; }

    // Copies of component pos / the whole state to the host.
    void read_state(std::size_t pos, RealType *host) const
    {
        check_hip(hipMemcpy(host, data_ + pos * n_, sizeof(RealType) * n_,
                      hipMemcpyDeviceToHost),
            "read_state");
    }

    // StateMatrix::print (state_matrix.hpp:172-188): one line per particle,
    // its components separated by sepchar (copies the state to the host).
    template <typename CharT, typename Traits>
    std::basic_ostream<CharT, Traits> &print(
        std::basic_ostream<CharT, Traits> &os, char sepchar = '\t') const
    {
        if (Dim == 0 || n_ == 0 || !os.good())
            return os;
        std::vector<RealType> h(Dim * n_);
        check(vsmch_synchronize(), "StateHIP::print");
        check_hip(hipMemcpy(h.data(), data_, sizeof(RealType) * h.size(),
                      hipMemcpyDeviceToHost),
            "StateHIP::print");
        for (size_type i = 0; i != n_; ++i) {
            for (std::size_t d = 0; d + 1 < Dim; ++d)
                os << h[d * n_ + i] << sepchar;
            os << h[(Dim - 1) * n_ + i] << '\n';
        }
        return os;
    }

    void bind_(RealType *p, vsmch_particle *h = nullptr)
    {
        data_ = p;
        h_ = h;
    }

  private:
    size_type n_;
    RealType *data_ = nullptr;
    vsmch_particle *h_ = nullptr;
    int scheme_ = VSMCH_RESAMPLE_STRATIFIED;

    void resample_event(size_type N)
    {
        if (N != n_)
            throw std::runtime_error("StateHIP::copy: N differs from the size");
        int r = 0;
        // the caller has decided: always resample
        check(vsmch_particle_resample(h_, scheme_,
                  std::numeric_limits<double>::infinity(), &r, nullptr, nullptr),
            "StateHIP::copy");
    }
};

template <typename CharT, typename Traits, std::size_t Dim, typename RealType>
inline std::basic_ostream<CharT, Traits> &operator<<(
    std::basic_ostream<CharT, Traits> &os, const StateHIP<Dim, RealType> &s)
{
    return s.print(os);
}

// SingleParticle<T> on the device: id, state access, the particle's engine.
template <typename T>
struct SingleParticleHIP {
    std::size_t id_;
    typename T::view_type v_;
    ParticleStream *rng_;

    __device__ std::size_t id() const { return id_; }
    __device__ typename T::state_type &state(std::size_t pos) const
    {
        return v_(id_, pos);
    }
    __device__ ParticleStream &rng() const { return *rng_; }
};

namespace detail {

// True for memory the device owns (hipMalloc); host arrays, pinned or not,
// are staged to the device first.
inline bool is_device_ptr(const void *p)
{
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// Device scratch of n doubles, grown on demand, freed with its owner.
class DeviceDoubles {
  public:
    DeviceDoubles() = default;
    DeviceDoubles(const DeviceDoubles &) = delete;
    DeviceDoubles &operator=(const DeviceDoubles &) = delete;
    ~DeviceDoubles()
    {
        if (p_)
            (void)hipFree(p_);
    }
    double *get(std::size_t n)
    {
        if (n > n_) {
            if (p_)
                check_hip(hipFree(p_), "scratch");
            p_ = nullptr;
            check_hip(hipMalloc(&p_, sizeof(double) * (n ? n : 1)), "scratch");
            n_ = n;
        }
        return p_;
    }

  private:
    double *p_ = nullptr;
    std::size_t n_ = 0;
};

} // namespace detail

// Weight (include/vsmc/core/weight.hpp:65-235) kept on the device, with the
// reference's members: set / mul / set_log / add_log (plain and with a
// stride), set_equal, ess, data, read_weight, draw. A pointer argument may
// be device or host memory (host arrays are copied to the device first);
// any other iterator is read on the host. The weights are kept as their
// logarithms: set and mul take non-negative linear weights.
// resample_data() is null: Particle::resample then hands the whole event to
// the state (particle.hpp:175-192), here StateHIP::copy(N, nullptr), which
// runs the library's device resampler.
class WeightHIP {
  public:
    using size_type = std::size_t;
    WeightHIP() = default;
    WeightHIP(vsmch_particle *h, size_type n)
        : h_(h), n_(n), buf_(std::make_shared<detail::DeviceDoubles>()),
          cum_(std::make_shared<detail::DeviceDoubles>())
    {
    }
    size_type size() const { return n_; }
    size_type resample_size() const { return n_; }
    double ess() const
    {
        double e = 0;
        check(vsmch_weight_ess(h_, &e), "Weight::ess");
        return e;
    }
    void set_equal() { check(vsmch_weight_set_equal(h_), "Weight::set_equal"); }

    template <typename It>
    void set(It first) { update(VSMCH_WEIGHT_SET, first, 1, "Weight::set"); }
    template <typename It>
    void set(It first, int stride) { update(VSMCH_WEIGHT_SET, first, stride, "Weight::set"); }
    template <typename It>
    void mul(It first) { update(VSMCH_WEIGHT_MUL, first, 1, "Weight::mul"); }
    template <typename It>
    void mul(It first, int stride) { update(VSMCH_WEIGHT_MUL, first, stride, "Weight::mul"); }
    template <typename It>
    void set_log(It first) { update(VSMCH_WEIGHT_SET_LOG, first, 1, "Weight::set_log"); }
    template <typename It>
    void set_log(It first, int stride)
    {
        update(VSMCH_WEIGHT_SET_LOG, first, stride, "Weight::set_log");
    }
    template <typename It>
    void add_log(It first) { update(VSMCH_WEIGHT_ADD_LOG, first, 1, "Weight::add_log"); }
    template <typename It>
    void add_log(It first, int stride)
    {
        update(VSMCH_WEIGHT_ADD_LOG, first, stride, "Weight::add_log");
    }

    const double *resample_data() const { return nullptr; }

    // Weight::data (weight.hpp:81): the normalised weights as a device
    // array, valid until the weights change
    const double *data() const
    {
        double *d = buf_->get(n_);
        check(vsmch_weight_read(h_, d), "Weight::data");
        return d;
    }

    // normalised weights to a host array or output iterator
    // (Weight::read_weight, weight.hpp:83-95)
    template <typename OutputIter>
    void read_weight(OutputIter first) const
    {
        const std::vector<double> w = host_weights();
        std::copy(w.begin(), w.end(), first);
    }
    template <typename RandomIter>
    void read_weight(RandomIter first, int stride) const
    {
        const std::vector<double> w = host_weights();
        for (size_type i = 0; i != n_; ++i, first += stride)
            *first = w[i];
    }
    void read_resample_weight(double *first) const { read_weight(first); }

    // Weight::draw (weight.hpp:175-179, DiscreteDistribution with
    // normalised weights, discrete_distribution.hpp:250-285): one U01 from
    // eng (U01Distribution<double>: u * 2^-32 from a 32-bit engine,
    // (u >> 11) * 2^-53 from a 64-bit one), then the first index whose
    // sequential cumulative weight is >= u (the last index if none). The
    // cumulative sums are the exact sequential ones, computed on the device.
    template <typename URNG>
    size_type draw(URNG &eng) const
    {
        static_assert(URNG::min() == 0, "draw: engine must start at 0");
        double u;
        if (static_cast<uint64_t>(URNG::max()) == 0xFFFFFFFFull)
            u = static_cast<double>(static_cast<uint32_t>(eng())) * 0x1p-32;
        else if (static_cast<uint64_t>(URNG::max()) == ~0ull)
            u = static_cast<double>(static_cast<uint64_t>(eng()) >> 11) * 0x1p-53;
        else
            throw std::runtime_error("Weight::draw: engine must give 32 or 64 bits");
        double *w = const_cast<double *>(data());
        double *c = cum_->get(n_);
        check(vsmch_exact_cumsum(n_, w, c), "Weight::draw");
        std::vector<double> h(n_);
        check(vsmch_synchronize(), "Weight::draw");
        check_hip(hipMemcpy(h.data(), c, sizeof(double) * n_, hipMemcpyDeviceToHost),
            "Weight::draw");
        const auto it = std::lower_bound(h.begin(), h.end(), u);
        return it == h.end() ? n_ - 1 : static_cast<size_type>(it - h.begin());
    }

    // log of the normalising constant accumulated by set_log / add_log
    // (user_guide/tex/smc.tex:169-182; the reference discards it)
    double log_zconst() const
    {
        double z = 0;
        check(vsmch_weight_log_zconst(h_, &z), "Weight::log_zconst");
        return z;
    }

  private:
    vsmch_particle *h_ = nullptr;
    size_type n_ = 0;
    std::shared_ptr<detail::DeviceDoubles> buf_;   // data(), staged inputs
    std::shared_ptr<detail::DeviceDoubles> cum_;   // draw()

    std::vector<double> host_weights() const
    {
        std::vector<double> h(n_);
        const double *d = data();
        check(vsmch_synchronize(), "Weight::read_weight");
        check_hip(hipMemcpy(h.data(), d, sizeof(double) * n_, hipMemcpyDeviceToHost),
            "Weight::read_weight");
        return h;
    }

    void update(int op, const double *first, int stride, const char *what)
    {
        if (stride < 1)
            throw std::runtime_error(std::string(what) + ": stride must be >= 1");
        if (detail::is_device_ptr(first)) {
            check(vsmch_weight_update(h_, op, first, stride), what);
            return;
        }
        std::vector<double> h(n_);
        for (size_type i = 0; i != n_; ++i)
            h[i] = first[i * static_cast<size_type>(stride)];
        stage(op, h, what);
    }
    void update(int op, double *first, int stride, const char *what)
    {
        update(op, const_cast<const double *>(first), stride, what);
    }
    template <typename It>
    void update(int op, It first, int stride, const char *what)
    {
        std::vector<double> h(n_);
        for (size_type i = 0; i != n_; ++i) {
            h[i] = static_cast<double>(*first);
            if (i + 1 != n_)
                std::advance(first, stride);
        }
        stage(op, h, what);
    }
    void stage(int op, const std::vector<double> &h, const char *what)
    {
        double *d = buf_->get(n_);
        check_hip(hipMemcpy(d, h.data(), sizeof(double) * n_, hipMemcpyHostToDevice), what);
        check(vsmch_weight_update(h_, op, d, 1), what);
        // the update reads the staging buffer on the library stream
        check(vsmch_synchronize(), what);
    }
};

// Particle<T> (include/vsmc/core/particle.hpp:47-210) over a device
// particle set. Engine of particle i: key seed_key(s, i), resampling engine
// seed_key(s, N), s the Seed state at construction.
template <typename T>
class Particle {
  public:
    using size_type = std::size_t;
    using value_type = T;
    using weight_type = WeightHIP;

    explicit Particle(size_type N) : size_(N), value_(N)
    {
        seed_ = Seed::instance().reserve(static_cast<uint64_t>(N) + 1);
        vsmch_particle_config cfg;
        cfg.n = static_cast<int64_t>(N);
        cfg.dim = static_cast<int>(T::dim());
        cfg.state_bytes = static_cast<int>(sizeof(typename T::state_type));
        cfg.seed = seed_;
        check(vsmch_particle_create(&h_, &cfg), "Particle");
        void *st = nullptr;
        check(vsmch_particle_state(h_, &st), "Particle");
        value_.bind_(static_cast<typename T::state_type *>(st), h_);
        weight_ = WeightHIP(h_, N);
        uint64_t *w = nullptr;
        check(vsmch_particle_rng_words(h_, &w), "Particle");
        words_ = reinterpret_cast<unsigned long long *>(w);
        check(vsmch_particle_scratch(h_, &scratch_), "Particle");
    }
    ~Particle()
    {
        if (h_)
            vsmch_particle_destroy(h_);
    }
    Particle(const Particle &) = delete;
    Particle &operator=(const Particle &) = delete;

    size_type size() const { return size_; }
    value_type &value() { return value_; }
    const value_type &value() const { return value_; }
    weight_type &weight() { return weight_; }
    const weight_type &weight() const { return weight_; }
    vsmch_particle *handle() const { return h_; }
    uint32_t seed() const { return seed_; }
    unsigned long long *rng_words() const { return words_; }
    // N doubles of device scratch a functor may use between eval_pre and
    // eval_post (the reference functors keep a Vector<double> w_ for this)
    double *scratch() const { return scratch_; }

    // Particle::resample(op, threshold) with op one of the schemes.
    bool resample(ResampleScheme scheme, double threshold)
    {
        int r = 0;
        check(vsmch_particle_resample(h_, static_cast<int>(scheme), threshold, &r,
                  nullptr, nullptr),
            "Particle::resample");
        return r != 0;
    }

  private:
    size_type size_;
    T value_;
    WeightHIP weight_;
    vsmch_particle *h_ = nullptr;
    uint32_t seed_ = 0;
    unsigned long long *words_ = nullptr;
    double *scratch_ = nullptr;
};

// ----------------------------------------------------------- functor bases

namespace detail {

__device__ inline void add_accept(unsigned long long *acc, unsigned long long v)
{
    // wave-level sum, one atomic per wave
    for (int d = 32; d >= 1; d >>= 1)
        v += __shfl_down(v, d, 64);
    if ((threadIdx.x % 64) == 0 && v)
        atomicAdd(acc, v);
}

template <typename T, typename D>
__global__ void init_kernel(D d, typename T::view_type v, uint32_t seed,
    unsigned long long *words, std::size_t n, unsigned long long *acc)
{
    const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    unsigned long long a = 0;
    if (i < n) {
        ParticleStream rng(seed_key(seed, i), words[i]);
        SingleParticleHIP<T> sp{i, v, &rng};
        a = static_cast<unsigned long long>(d.eval_sp(sp));
        words[i] = rng.word;
    }
    add_accept(acc, a);
}

template <typename T, typename D>
__global__ void move_kernel(D d, std::size_t iter, typename T::view_type v,
    uint32_t seed, unsigned long long *words, std::size_t n,
    unsigned long long *acc)
{
    const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    unsigned long long a = 0;
    if (i < n) {
        ParticleStream rng(seed_key(seed, i), words[i]);
        SingleParticleHIP<T> sp{i, v, &rng};
        a = static_cast<unsigned long long>(d.eval_sp(iter, sp));
        words[i] = rng.word;
    }
    add_accept(acc, a);
}

template <typename T, typename D>
__global__ void monitor_kernel(D d, std::size_t iter, std::size_t dim,
    typename T::view_type v, uint32_t seed, unsigned long long *words,
    std::size_t n, double *r)
{
    const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        ParticleStream rng(seed_key(seed, i), words[i]);
        SingleParticleHIP<T> sp{i, v, &rng};
        d.eval_sp(iter, dim, sp, r + i * dim);
        words[i] = rng.word;
    }
}

class AcceptCounter {
  public:
    // one counter for the process; not freed at exit (the HIP runtime may
    // already be torn down when static destructors run)
    AcceptCounter() { check_hip(hipMalloc(&p_, sizeof(unsigned long long)), "accept"); }
    unsigned long long *reset()
    {
        check_hip(hipMemset(p_, 0, sizeof(unsigned long long)), "accept");
        return p_;
    }
    std::size_t read()
    {
        unsigned long long h = 0;
        check(vsmch_synchronize(), "accept");
        check_hip(hipMemcpy(&h, p_, sizeof(h), hipMemcpyDeviceToHost), "accept");
        return static_cast<std::size_t>(h);
    }

  private:
    unsigned long long *p_ = nullptr;
};

inline AcceptCounter &accept_counter()
{
    static AcceptCounter c;
    return c;
}

inline dim3 grid_for(std::size_t n)
{
    return dim3(static_cast<unsigned int>((n + 255) / 256));
}

} // namespace detail

// InitializeSMP shape (backend_omp.hpp:64-81): eval_param, eval_pre,
// eval_sp over all particles (device), eval_post; returns the accept count.
template <typename T, typename Derived>
class InitializeHIP {
  public:
    std::size_t operator()(Particle<T> &particle, void *param)
    {
        Derived &d = static_cast<Derived &>(*this);
        d.eval_param(particle, param);
        d.eval_pre(particle);
        unsigned long long *acc = detail::accept_counter().reset();
        hipLaunchKernelGGL((detail::init_kernel<T, Derived>),
            detail::grid_for(particle.size()), dim3(256), 0, nullptr, d,
            particle.value().view(), particle.seed(), particle.rng_words(),
            particle.size(), acc);
        check_hip(hipGetLastError(), "InitializeHIP");
        d.eval_post(particle);
        return detail::accept_counter().read();
    }
    void eval_param(Particle<T> &, void *) {}
    void eval_pre(Particle<T> &) {}
    void eval_post(Particle<T> &) {}
};

// MoveSMP shape (backend_omp.hpp:89-105)
template <typename T, typename Derived>
class MoveHIP {
  public:
    std::size_t operator()(std::size_t iter, Particle<T> &particle)
    {
        Derived &d = static_cast<Derived &>(*this);
        d.eval_pre(iter, particle);
        unsigned long long *acc = detail::accept_counter().reset();
        hipLaunchKernelGGL((detail::move_kernel<T, Derived>),
            detail::grid_for(particle.size()), dim3(256), 0, nullptr, d, iter,
            particle.value().view(), particle.seed(), particle.rng_words(),
            particle.size(), acc);
        check_hip(hipGetLastError(), "MoveHIP");
        d.eval_post(iter, particle);
        return detail::accept_counter().read();
    }
    void eval_pre(std::size_t, Particle<T> &) {}
    void eval_post(std::size_t, Particle<T> &) {}
};

// MonitorEvalSMP shape (backend_omp.hpp:113-129): eval_sp writes dim values
// per particle into r (device, N x dim row-major).
template <typename T, typename Derived>
class MonitorEvalHIP {
  public:
    void operator()(std::size_t iter, std::size_t dim, Particle<T> &particle,
        double *r)
    {
        Derived &d = static_cast<Derived &>(*this);
        d.eval_pre(iter, particle);
        hipLaunchKernelGGL((detail::monitor_kernel<T, Derived>),
            detail::grid_for(particle.size()), dim3(256), 0, nullptr, d, iter,
            dim, particle.value().view(), particle.seed(), particle.rng_words(),
            particle.size(), r);
        check_hip(hipGetLastError(), "MonitorEvalHIP");
        d.eval_post(iter, particle);
    }
    void eval_pre(std::size_t, Particle<T> &) {}
    void eval_post(std::size_t, Particle<T> &) {}
};

// Monitor<T> (include/vsmc/core/monitor.hpp:60-304). The eval functor fills
// an N x dim device buffer; the weighted sum r_d = sum_i W_i buf[i][d] (the
// reference's dgemv, :263-266) runs in the library. record_only monitors
// get a host double[dim] directly (:253-257).
template <typename T>
class Monitor {
  public:
    using eval_type =
        std::function<void(std::size_t, std::size_t, Particle<T> &, double *)>;

    Monitor(std::size_t dim, const eval_type &eval, bool record_only = false,
        MonitorStage stage = MonitorMCMC)
        : dim_(dim), eval_(eval), record_only_(record_only), stage_(stage),
          name_(dim)
    {
    }
    Monitor(const Monitor &o)
        : dim_(o.dim_), eval_(o.eval_), recording_(o.recording_),
          record_only_(o.record_only_), stage_(o.stage_), name_(o.name_),
          index_(o.index_), record_(o.record_)
    {
    }
    ~Monitor()
    {
        if (buf_)
            (void)hipFree(buf_);
    }

    std::size_t dim() const { return dim_; }
    bool record_only() const { return record_only_; }
    MonitorStage stage() const { return stage_; }
    std::size_t iter_size() const { return index_.size(); }
    bool empty() const { return !static_cast<bool>(eval_); }
    std::string &name(std::size_t id) { return name_[id]; }
    const std::string &name(std::size_t id) const { return name_[id]; }
    std::size_t index(std::size_t iter) const { return index_[iter]; }
    double record(std::size_t id, std::size_t iter) const
    {
        return record_[iter * dim_ + id];
    }
    const double *record_data(std::size_t iter) const
    {
        return record_.data() + iter * dim_;
    }
    void clear()
    {
        index_.clear();
        record_.clear();
    }
    bool recording() const { return recording_; }
    void turn_on() { recording_ = true; }
    void turn_off() { recording_ = false; }

    void eval(std::size_t iter, Particle<T> &particle, MonitorStage stage)
    {
        if (!recording_ || stage != stage_)
            return;
        std::vector<double> result(dim_);
        if (record_only_) {
            eval_(iter, dim_, particle, result.data());
        } else {
            const std::size_t need = particle.size() * dim_;
            if (need > cap_) {
                if (buf_)
                    check_hip(hipFree(buf_), "Monitor");
                check_hip(hipMalloc(&buf_, sizeof(double) * need), "Monitor");
                cap_ = need;
            }
            eval_(iter, dim_, particle, buf_);
            check(vsmch_weighted_sum(particle.handle(), static_cast<int>(dim_),
                      buf_, result.data()),
                "Monitor::eval");
        }
        index_.push_back(iter);
        record_.insert(record_.end(), result.begin(), result.end());
    }

  private:
    std::size_t dim_;
    eval_type eval_;
    bool recording_ = true;
    bool record_only_;
    MonitorStage stage_;
    std::vector<std::string> name_;
    std::vector<std::size_t> index_;
    std::vector<double> record_;
    double *buf_ = nullptr;
    std::size_t cap_ = 0;
};

// Sampler<T> (include/vsmc/core/sampler.hpp:57-869): the step order, the
// histories and the summary output of the reference.
template <typename T>
class Sampler {
  public:
    using size_type = std::size_t;
    using value_type = T;
    using init_type = std::function<std::size_t(Particle<T> &, void *)>;
    using move_type = std::function<std::size_t(std::size_t, Particle<T> &)>;
    using mcmc_type = std::function<std::size_t(std::size_t, Particle<T> &)>;
    using monitor_map_type = std::map<std::string, Monitor<T>>;

    explicit Sampler(size_type N)
        : particle_(N), scheme_(Multinomial), threshold_(resample_threshold_never())
    {
    }
    Sampler(size_type N, ResampleScheme scheme)
        : particle_(N), scheme_(scheme), threshold_(resample_threshold_always())
    {
    }
    Sampler(size_type N, ResampleScheme scheme, double threshold)
        : particle_(N), scheme_(scheme), threshold_(threshold)
    {
    }

    size_type size() const { return particle_.size(); }
    std::size_t iter_size() const { return size_history_.size(); }
    std::size_t iter_num() const { return iter_num_; }
    Particle<T> &particle() { return particle_; }
    const Particle<T> &particle() const { return particle_; }

    Sampler &resample_scheme(ResampleScheme scheme)
    {
        scheme_ = scheme;
        return *this;
    }
    double resample_threshold() const { return threshold_; }
    Sampler &resample_threshold(double t)
    {
        threshold_ = t;
        return *this;
    }
    static double resample_threshold_never()
    {
        return -std::numeric_limits<double>::infinity();
    }
    static double resample_threshold_always()
    {
        return std::numeric_limits<double>::infinity();
    }
    Sampler &resample()
    {
        particle_.resample(scheme_, resample_threshold_always());
        return *this;
    }

    Sampler &init(const init_type &f)
    {
        init_ = f;
        return *this;
    }
    Sampler &move(const move_type &f, bool append)
    {
        if (!append)
            move_queue_.clear();
        move_queue_.push_back(f);
        return *this;
    }
    Sampler &mcmc(const mcmc_type &f, bool append)
    {
        if (!append)
            mcmc_queue_.clear();
        mcmc_queue_.push_back(f);
        return *this;
    }
    Sampler &monitor(const std::string &name, std::size_t dim,
        const typename Monitor<T>::eval_type &eval, bool record_only = false,
        MonitorStage stage = MonitorMCMC)
    {
        monitor_.insert(std::make_pair(name, Monitor<T>(dim, eval, record_only, stage)));
        return *this;
    }
    Monitor<T> &monitor(const std::string &name)
    {
        auto it = monitor_.find(name);
        if (it == monitor_.end())
            throw RuntimeError("Sampler::monitor: no monitor named " + name);
        return it->second;
    }

    double size_history(std::size_t iter) const { return size_history_[iter]; }
    double ess_history(std::size_t iter) const { return ess_history_[iter]; }
    bool resampled_history(std::size_t iter) const { return resampled_history_[iter]; }
    std::size_t accept_history(std::size_t id, std::size_t iter) const
    {
        return accept_history_[id][iter];
    }
    // log-evidence after each step (an extension; smc.tex:169-182)
    double log_zconst_history(std::size_t iter) const { return logz_history_[iter]; }

    // sampler.hpp:443-456
    Sampler &initialize(void *param = nullptr)
    {
        do_reset();
        do_acch();
        if (!init_)
            throw RuntimeError("Sampler::initialize: no init functor");
        accept_history_[0].push_back(init_(particle_, param));
        after_move(1);
        do_acch();
        return *this;
    }

    // sampler.hpp:464-476
    Sampler &iterate(std::size_t num = 1)
    {
        do_acch();
        for (std::size_t i = 0; i != num; ++i) {
            ++iter_num_;
            std::size_t ia = 0;
            for (auto &m : move_queue_)
                accept_history_[ia++].push_back(m(iter_num_, particle_));
            after_move(ia);
        }
        do_acch();
        return *this;
    }

    // Summary header and data (sampler.hpp:551-648, :786-865): integer
    // columns Size, Resampled, Accept.k; floating-point columns ESS and each
    // recording monitor's components (NaN where a monitor did not record).
    std::size_t summary_header_size_int() const
    {
        return iter_size() == 0 ? 0 : accept_history_.size() + 2;
    }
    std::size_t summary_header_size() const
    {
        if (iter_size() == 0)
            return 0;
        std::size_t n = 1;
        for (const auto &m : monitor_)
            if (m.second.iter_size() > 0)
                n += m.second.dim();
        return n;
    }
    template <typename OutputIter>
    void summary_header_int(OutputIter first) const
    {
        if (summary_header_size_int() == 0)
            return;
        *first++ = std::string("Size");
        *first++ = std::string("Resampled");
        for (std::size_t i = 0; i != accept_history_.size(); ++i)
            *first++ = "Accept." + std::to_string(i);
    }
    template <typename OutputIter>
    void summary_header(OutputIter first) const
    {
        if (summary_header_size() == 0)
            return;
        *first++ = std::string("ESS");
        for (const auto &m : monitor_) {
            if (m.second.iter_size() == 0)
                continue;
            for (std::size_t d = 0; d != m.second.dim(); ++d)
                *first++ = m.second.name(d).empty()
                    ? m.first + "." + std::to_string(d)
                    : m.second.name(d);
        }
    }
    std::size_t summary_data_size_int() const
    {
        return summary_header_size_int() * iter_size();
    }
    std::size_t summary_data_size() const
    {
        return summary_header_size() * iter_size();
    }
    template <MatrixLayout Layout, typename OutputIter>
    void summary_data_int(OutputIter first) const
    {
        if (summary_data_size_int() == 0)
            return;
        using V = typename std::iterator_traits<OutputIter>::value_type;
        const std::size_t nc = summary_header_size_int();
        for (std::size_t r = 0; r != iter_size(); ++r) {
            const auto put = [&](std::size_t c, std::size_t v) {
                const std::size_t at = Layout == RowMajor ? r * nc + c : c * iter_size() + r;
                first[at] = static_cast<V>(v);
            };
            put(0, size_history_[r]);
            put(1, resampled_history_[r] ? 1 : 0);
            for (std::size_t i = 0; i != accept_history_.size(); ++i)
                put(2 + i, accept_history_[i][r]);
        }
    }
    template <MatrixLayout Layout, typename OutputIter>
    void summary_data(OutputIter first) const
    {
        if (summary_data_size() == 0)
            return;
        const std::size_t nc = summary_header_size();
        const std::size_t nr = iter_size();
        const auto at = [&](std::size_t r, std::size_t c) {
            return Layout == RowMajor ? r * nc + c : c * nr + r;
        };
        for (std::size_t r = 0; r != nr; ++r)
            first[at(r, 0)] = ess_history_[r];
        std::size_t c0 = 1;
        for (const auto &m : monitor_) {
            if (m.second.iter_size() == 0)
                continue;
            const std::size_t md = m.second.dim();
            std::size_t k = 0;
            for (std::size_t r = 0; r != nr; ++r) {
                // the r-th row holds the k-th record when it was taken at r
                const bool has = k != m.second.iter_size() && r == m.second.index(k);
                for (std::size_t d = 0; d != md; ++d)
                    first[at(r, c0 + d)] = has ? m.second.record(d, k)
                                               : std::numeric_limits<double>::quiet_NaN();
                k += has ? 1 : 0;
            }
            c0 += md;
        }
    }

    // sampler.hpp:651-689: tab-separated Size, Resampled, Accept.k, ESS,
    // then each monitor's columns.
    template <typename CharT, typename Traits>
    std::basic_ostream<CharT, Traits> &print(
        std::basic_ostream<CharT, Traits> &os, char sep = '\t') const
    {
        if (iter_size() == 0 || !os.good())
            return os;
        os << "Size" << sep << "Resampled" << sep;
        for (std::size_t i = 0; i != accept_history_.size(); ++i)
            os << "Accept." << i << sep;
        os << "ESS" << sep;
        for (const auto &m : monitor_) {
            if (m.second.iter_size() == 0)
                continue;
            for (std::size_t d = 0; d != m.second.dim(); ++d) {
                if (m.second.name(d).empty())
                    os << m.first << "." << d << sep;
                else
                    os << m.second.name(d) << sep;
            }
        }
        os << '\n';
        std::vector<std::size_t> mi(monitor_.size(), 0);
        for (std::size_t r = 0; r != iter_size(); ++r) {
            os << size_history_[r] << sep << resampled_history_[r] << sep;
            for (std::size_t i = 0; i != accept_history_.size(); ++i)
                os << accept_history_[i][r] << sep;
            os << ess_history_[r] << sep;
            std::size_t k = 0;
            for (const auto &m : monitor_) {
                if (m.second.iter_size() == 0) {
                    ++k;
                    continue;
                }
                if (mi[k] != m.second.iter_size() && r == m.second.index(mi[k])) {
                    for (std::size_t d = 0; d != m.second.dim(); ++d)
                        os << m.second.record(d, mi[k]) << sep;
                    ++mi[k];
                } else {
                    for (std::size_t d = 0; d != m.second.dim(); ++d)
                        os << std::numeric_limits<double>::quiet_NaN() << sep;
                }
                ++k;
            }
            os << '\n';
        }
        return os;
    }

  private:
    Particle<T> particle_;
    ResampleScheme scheme_;
    double threshold_;
    std::size_t iter_num_ = 0;
    init_type init_;
    std::vector<move_type> move_queue_;
    std::vector<mcmc_type> mcmc_queue_;
    std::vector<std::size_t> size_history_;
    std::vector<double> ess_history_;
    std::vector<bool> resampled_history_;
    std::vector<double> logz_history_;
    std::vector<std::vector<std::size_t>> accept_history_;
    monitor_map_type monitor_;

    // monitor(Move), resample, monitor(Resample), MCMC, monitor(MCMC)
    // (sampler.hpp:746-754, :772-778)
    void after_move(std::size_t ia)
    {
        do_monitor(MonitorMove);
        size_history_.push_back(size());
        ess_history_.push_back(particle_.weight().ess());
        logz_history_.push_back(particle_.weight().log_zconst());
        resampled_history_.push_back(particle_.resample(scheme_, threshold_));
        do_monitor(MonitorResample);
        for (auto &m : mcmc_queue_)
            accept_history_[ia++].push_back(m(iter_num_, particle_));
        do_monitor(MonitorMCMC);
    }

    void do_monitor(MonitorStage stage)
    {
        for (auto &m : monitor_)
            if (!m.second.empty())
                m.second.eval(iter_num_, particle_, stage);
    }

    void do_acch()
    {
        if (accept_history_.empty())
            accept_history_.push_back(std::vector<std::size_t>());
        // initialize() records the init functor in column 0 and its MCMC
        // functors from column 1 (sampler.hpp:735-744), so with no move
        // functor one more column than moves + mcmc is needed; the
        // reference sizes to moves + mcmc and writes past the end there.
        const std::size_t nmove = move_queue_.empty() ? 1 : move_queue_.size();
        const std::size_t n = nmove + mcmc_queue_.size();
        while (accept_history_.size() < n)
            accept_history_.push_back(std::vector<std::size_t>());
        for (auto &a : accept_history_)
            a.resize(iter_size());
    }

    void do_reset()
    {
        size_history_.clear();
        ess_history_.clear();
        resampled_history_.clear();
        logz_history_.clear();
        accept_history_.clear();
        for (auto &m : monitor_)
            m.second.clear();
        iter_num_ = 0;
        particle_.weight().set_equal();
    }
};

template <typename CharT, typename Traits, typename T>
inline std::basic_ostream<CharT, Traits> &operator<<(
    std::basic_ostream<CharT, Traits> &os, const Sampler<T> &s)
{
    return s.print(os);
}

} // namespace vsmch
