// The reference's own plug points over the HIP layer.
//
// Compile time: the members that Particle<T>::resample
// (include/vsmc/core/particle.hpp:170-197), the Weight interface
// (include/vsmc/core/weight.hpp:74-193) and a user weight_type
// (user_guide/tex/advanced.tex:67-73) require, checked on WeightHIP and
// StateHIP with static_asserts.
//
// Run time (args: N scheme input-file): reads a state (3 x N doubles, hex)
// and log-weights (N doubles, hex), sets them through the host-pointer forms,
// then drives one event with the control flow of Particle::resample
// restated below (decision from ess(); resample_data() null, so the whole
// event goes to copy(N, nullptr); then set_equal()), and exercises copy with
// a host index array, the strided set / mul / set_log / add_log forms,
// data(), read_weight and draw with 32- and 64-bit engines. Everything is
// printed in hex for the test to check against the oracle.
#include <vsmch/sampler.hpp>

#include <cstdio>
#include <fstream>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

namespace {

using T = vsmch::StateHIP<3, double>;

// --- compile-time conformance
template <typename W>
using ess_r = decltype(std::declval<const W &>().ess());
template <typename W>
using rsize_r = decltype(std::declval<const W &>().resample_size());
template <typename W>
using rdata_r = decltype(std::declval<const W &>().resample_data());
template <typename W>
using data_r = decltype(std::declval<const W &>().data());
template <typename W>
using draw_r = decltype(std::declval<const W &>().draw(std::declval<std::mt19937 &>()));

static_assert(std::is_convertible<ess_r<vsmch::WeightHIP>, double>::value, "ess()");
static_assert(std::is_convertible<rsize_r<vsmch::WeightHIP>, std::size_t>::value,
    "resample_size()");
static_assert(std::is_convertible<rdata_r<vsmch::WeightHIP>, const double *>::value,
    "resample_data()");
static_assert(std::is_convertible<data_r<vsmch::WeightHIP>, const double *>::value, "data()");
static_assert(std::is_convertible<draw_r<vsmch::WeightHIP>, std::size_t>::value, "draw()");
static_assert(std::is_same<decltype(std::declval<vsmch::WeightHIP &>().set_equal()), void>::value,
    "set_equal()");
template <typename W, typename P>
using set_r = decltype(std::declval<W &>().set(std::declval<P>()),
    std::declval<W &>().set(std::declval<P>(), 2), std::declval<W &>().mul(std::declval<P>()),
    std::declval<W &>().mul(std::declval<P>(), 2), std::declval<W &>().set_log(std::declval<P>()),
    std::declval<W &>().set_log(std::declval<P>(), 2),
    std::declval<W &>().add_log(std::declval<P>()),
    std::declval<W &>().add_log(std::declval<P>(), 2),
    std::declval<const W &>().read_weight(std::declval<double *>()),
    std::declval<const W &>().read_weight(std::declval<double *>(), 2), 0);
static_assert(std::is_same<set_r<vsmch::WeightHIP, const double *>, int>::value, "set/mul forms");
static_assert(std::is_same<set_r<vsmch::WeightHIP, std::vector<double>::iterator>, int>::value,
    "iterator forms");
template <typename S>
using copy_r = decltype(std::declval<S &>().copy(std::size_t(), std::declval<const std::size_t *>()),
    std::declval<S &>().copy(std::size_t(), std::declval<const int *>()),
    std::declval<S &>().copy(std::size_t(), nullptr), 0);
static_assert(std::is_same<copy_r<T>, int>::value, "copy(N, index) and copy(N, nullptr)");
static_assert(std::is_constructible<T, std::size_t>::value, "T(N)");
static_assert(T::dim() == 3, "dim()");

// --- Particle::resample's control flow (particle.hpp:170-197), restated
template <typename P>
bool reference_resample(P &particle, double threshold)
{
    const std::size_t N = particle.weight().resample_size();
    const bool resampled = particle.weight().ess() < threshold * N;
    if (resampled) {
        const double *rw = particle.weight().resample_data();
        if (rw != nullptr)
            throw std::runtime_error("host weights were not expected");
        particle.value().copy(N, static_cast<const std::size_t *>(nullptr));
        particle.weight().set_equal();
    }
    return resampled;
}

double rd(std::ifstream &in)
{
    std::string t;
    in >> t;
    return std::stod(t);   // accepts hex floating point
}

void print_vec(const char *tag, const std::vector<double> &v)
{
    std::printf("%s", tag);
    for (double x : v)
        std::printf("\t%a", x);
    std::printf("\n");
}

std::vector<double> state_of(vsmch::Particle<T> &p)
{
    std::vector<double> h(3 * p.size());
    vsmch::check(vsmch_synchronize(), "state");
    vsmch::check_hip(hipMemcpy(h.data(), p.value().data(), sizeof(double) * h.size(),
                         hipMemcpyDeviceToHost),
        "state");
    return h;
}

std::vector<double> weights_of(vsmch::Particle<T> &p)
{
    std::vector<double> w(p.size());
    p.weight().read_weight(w.data());
    return w;
}

} // namespace

int main(int argc, char **argv)
{
    if (argc != 4) {
        std::fprintf(stderr, "usage: plug_hip N scheme input\n");
        return 2;
    }
    const std::size_t N = std::stoul(argv[1]);
    const int scheme = std::stoi(argv[2]);
    std::ifstream in(argv[3]);
    std::vector<double> x(3 * N), lw(N), w2(2 * N);
    std::vector<std::size_t> idx(N);
    for (auto &v : x)
        v = rd(in);
    for (auto &v : lw)
        v = rd(in);
    for (auto &v : idx) {
        std::string t;
        in >> t;
        v = std::stoul(t);
    }
    for (auto &v : w2)
        v = rd(in);

    vsmch::Particle<T> p(N);
    p.value().resample_scheme(scheme);
    std::printf("SEED\t%u\n", p.seed());
    vsmch::check_hip(hipMemcpy(p.value().data(), x.data(), sizeof(double) * x.size(),
                         hipMemcpyHostToDevice),
        "state");
    p.weight().set_log(lw.data());   // a host pointer
    print_vec("W0", weights_of(p));
    std::printf("ESS0\t%a\n", p.weight().ess());
    const bool r = reference_resample(p, 0.5);
    std::printf("RESAMPLED\t%d\n", r ? 1 : 0);
    print_vec("X1", state_of(p));
    print_vec("W1", weights_of(p));
    // copy with a host index array
    p.value().copy(N, idx.data());
    print_vec("X2", state_of(p));
    // strided linear weights from the host, then a device multiplication
    p.weight().set(w2.data(), 2);
    print_vec("WSET", weights_of(p));
    p.weight().mul(w2.data() + 1, 2);
    print_vec("WMUL", weights_of(p));
    // log forms: set_log through an iterator, add_log from a device array
    // with a stride
    p.weight().set_log(lw.begin(), 1);
    print_vec("WSETLOG", weights_of(p));
    double *dinc = nullptr;
    vsmch::check_hip(hipMalloc(&dinc, sizeof(double) * 2 * N), "inc");
    vsmch::check_hip(hipMemcpy(dinc, w2.data(), sizeof(double) * 2 * N, hipMemcpyHostToDevice),
        "inc");
    p.weight().add_log(static_cast<const double *>(dinc), 2);
    print_vec("WADDLOG", weights_of(p));
    vsmch::check_hip(hipFree(dinc), "inc");
    std::vector<double> h(N);
    vsmch::check(vsmch_synchronize(), "data");
    vsmch::check_hip(hipMemcpy(h.data(), p.weight().data(), sizeof(double) * N,
                         hipMemcpyDeviceToHost),
        "data");
    print_vec("DATA", h);
    std::vector<double> wr(2 * N, -1.0);
    p.weight().read_weight(wr.data(), 2);
    print_vec("READ2", wr);
    // draws: the engine's word and the drawn index, 32- and 64-bit engines
    std::mt19937 e32(5);
    std::mt19937 c32(5);
    for (int k = 0; k < 8; ++k) {
        const unsigned long long word = c32();
        std::printf("DRAW32\t%llu\t%zu\n", word, p.weight().draw(e32));
    }
    std::mt19937_64 e64(7);
    std::mt19937_64 c64(7);
    for (int k = 0; k < 8; ++k) {
        const unsigned long long word = c64();
        std::printf("DRAW64\t%llu\t%zu\n", word, p.weight().draw(e64));
    }
    return 0;
}
