// pf_hip.cpp -- the constant-velocity PF of user_guide/cpp/pf_seq.cpp
// written against the vsmch C++ layer, with the reference's structure: a
// state class, PFInit / PFMove / PFMEval functors with eval_param / eval_pre
// / eval_sp / eval_post, a Sampler with a "pos" monitor, initialize +
// iterate, and the summary printed with operator<<. The eval_sp bodies are
// __device__; everything else is host code as in the reference.
//
// Build (done by __graft_entry__.build()):
//   hipcc -O3 -ffp-contract=off --offload-arch=gfx950 -Iinclude \
//     tests/cpp/pf_hip.cpp -Lvsmc_amd -lvsmch -o tests/cpp/pf_hip
// Usage: pf_hip N T scheme threshold seed obs_file [full]
// With "full", also prints the summary header / data in both layouts
// (lines SHI, SDIR, SDIC, SH, SDR, SDC) and the final state (STATE, then N
// lines from the state's operator<<).
#include <vsmch/sampler.hpp>

#include <chrono>

#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

static const std::size_t PosX = 0;
static const std::size_t PosY = 1;
static const std::size_t VelX = 2;
static const std::size_t VelY = 3;

// example/pf/include/pf.hpp:78-87 (scale 1/0.1, nu = 10)
__device__ inline double pf_log_likelihood(double x, double y, double ox,
    double oy)
{
    const double scale = 10;
    const double nu = 10;
    double llh_x = scale * (x - ox);
    double llh_y = scale * (y - oy);
    llh_x = log(1 + llh_x * llh_x / nu);
    llh_y = log(1 + llh_y * llh_y / nu);
    return -0.5 * (nu + 1) * (llh_x + llh_y);
}

class PFState : public vsmch::StateHIP<4, double> {
  public:
    explicit PFState(size_type N) : vsmch::StateHIP<4, double>(N) {}
    ~PFState()
    {
        if (obs_)
            (void)hipFree(obs_);
    }

    void read_data(const char *file)
    {
        std::vector<double> x, y;
        std::ifstream data(file);
        double a, b;
        while (data >> a >> b) {
            x.push_back(a);
            y.push_back(b);
        }
        if (x.empty())
            throw vsmch::RuntimeError(std::string("no observations in ") + file);
        n_obs_ = x.size();
        if (obs_)
            (void)hipFree(obs_);
        vsmch::check_hip(hipMalloc(&obs_, sizeof(double) * 2 * n_obs_), "obs");
        vsmch::check_hip(hipMemcpy(obs_, x.data(), sizeof(double) * n_obs_,
                             hipMemcpyHostToDevice),
            "obs");
        vsmch::check_hip(hipMemcpy(obs_ + n_obs_, y.data(),
                             sizeof(double) * n_obs_, hipMemcpyHostToDevice),
            "obs");
    }
    const double *obs_x() const { return obs_; }
    const double *obs_y() const { return obs_ + n_obs_; }

  private:
    double *obs_ = nullptr;
    std::size_t n_obs_ = 0;
};

class PFInit : public vsmch::InitializeHIP<PFState, PFInit> {
  public:
    void eval_param(vsmch::Particle<PFState> &particle, void *param)
    {
        particle.value().read_data(static_cast<const char *>(param));
    }
    void eval_pre(vsmch::Particle<PFState> &particle)
    {
        w_ = particle.scratch();
        ox_ = particle.value().obs_x();
        oy_ = particle.value().obs_y();
    }
    __device__ std::size_t eval_sp(vsmch::SingleParticleHIP<PFState> sp)
    {
        vsmch::NormalDistribution<double> norm_pos(0, 2);
        vsmch::NormalDistribution<double> norm_vel(0, 1);
        sp.state(PosX) = norm_pos(sp.rng());
        sp.state(PosY) = norm_pos(sp.rng());
        sp.state(VelX) = norm_vel(sp.rng());
        sp.state(VelY) = norm_vel(sp.rng());
        w_[sp.id()] = pf_log_likelihood(sp.state(PosX), sp.state(PosY), ox_[0],
            oy_[0]);
        return 0;
    }
    void eval_post(vsmch::Particle<PFState> &particle)
    {
        particle.weight().set_log(w_);
    }

  private:
    double *w_ = nullptr;
    const double *ox_ = nullptr;
    const double *oy_ = nullptr;
};

class PFMove : public vsmch::MoveHIP<PFState, PFMove> {
  public:
    void eval_pre(std::size_t, vsmch::Particle<PFState> &particle)
    {
        w_ = particle.scratch();
        ox_ = particle.value().obs_x();
        oy_ = particle.value().obs_y();
    }
    __device__ std::size_t eval_sp(std::size_t t,
        vsmch::SingleParticleHIP<PFState> sp)
    {
        vsmch::NormalDistribution<double> norm_pos(0, 0.02);
        vsmch::NormalDistribution<double> norm_vel(0, 0.02);
        sp.state(PosX) += norm_pos(sp.rng()) + 0.1 * sp.state(VelX);
        sp.state(PosY) += norm_pos(sp.rng()) + 0.1 * sp.state(VelY);
        sp.state(VelX) += norm_vel(sp.rng());
        sp.state(VelY) += norm_vel(sp.rng());
        w_[sp.id()] = pf_log_likelihood(sp.state(PosX), sp.state(PosY), ox_[t],
            oy_[t]);
        return 0;
    }
    void eval_post(std::size_t, vsmch::Particle<PFState> &particle)
    {
        particle.weight().add_log(w_);
    }

  private:
    double *w_ = nullptr;
    const double *ox_ = nullptr;
    const double *oy_ = nullptr;
};

class PFMEval : public vsmch::MonitorEvalHIP<PFState, PFMEval> {
  public:
    __device__ void eval_sp(std::size_t, std::size_t,
        vsmch::SingleParticleHIP<PFState> sp, double *r)
    {
        r[0] = sp.state(PosX);
        r[1] = sp.state(PosY);
    }
};

int main(int argc, char **argv)
{
    if (argc < 7) {
        std::cerr << "usage: pf_hip N T scheme threshold seed obs_file\n";
        return 2;
    }
    const std::size_t N = std::strtoull(argv[1], nullptr, 10);
    const std::size_t T = std::strtoull(argv[2], nullptr, 10);
    const int scheme = std::atoi(argv[3]);
    const double threshold = std::atof(argv[4]);
    const unsigned seed = static_cast<unsigned>(std::strtoul(argv[5], nullptr, 10));
    {
        // the move reads observation t for t < T: check the file first
        std::ifstream data(argv[6]);
        double a, b;
        std::size_t n = 0;
        while (data >> a >> b)
            ++n;
        if (n < T || T == 0) {
            std::cerr << "pf_hip: " << argv[6] << " holds " << n
                      << " observation pairs, " << T << " needed\n";
            return 2;
        }
    }
    try {
        vsmch::Seed::instance().set(seed);
        vsmch::Sampler<PFState> sampler(N,
            static_cast<vsmch::ResampleScheme>(scheme), threshold);
        sampler.init(PFInit()).move(PFMove(), false).monitor("pos", 2, PFMEval());
        if (argc > 7 && std::string(argv[7]) == "time") {
            // initialize + 3 untimed steps, then T - 4 timed steps (host
            // clock around the steps, the queue drained on both sides)
            sampler.initialize(argv[6]).iterate(3);
            vsmch::check(vsmch_synchronize(), "time");
            vsmch::check_hip(hipDeviceSynchronize(), "time");
            const auto t0 = std::chrono::steady_clock::now();
            sampler.iterate(T - 4);
            vsmch::check(vsmch_synchronize(), "time");
            vsmch::check_hip(hipDeviceSynchronize(), "time");
            const auto t1 = std::chrono::steady_clock::now();
            std::size_t nres = 0;
            for (std::size_t t = 4; t != T; ++t)
                nres += sampler.resampled_history(t) ? 1 : 0;
            std::cout << "TIME\t" << (T - 4) << '\t'
                      << std::chrono::duration<double, std::milli>(t1 - t0).count()
                      << '\t' << nres << '\n';
            return 0;
        }
        sampler.initialize(argv[6]).iterate(T - 1);
        std::cout.precision(17);
        std::cout << sampler;
        std::cout << "LOGZ";
        for (std::size_t t = 0; t != T; ++t)
            std::cout << '\t' << sampler.log_zconst_history(t);
        std::cout << '\n';
        if (argc > 7 && std::string(argv[7]) == "full") {
            std::vector<std::string> hi(sampler.summary_header_size_int());
            std::vector<std::string> hf(sampler.summary_header_size());
            sampler.summary_header_int(hi.begin());
            sampler.summary_header(hf.begin());
            std::vector<long long> di(sampler.summary_data_size_int());
            std::vector<double> df(sampler.summary_data_size());
            const auto line = [](const char *tag, const auto &v) {
                std::cout << tag;
                for (const auto &x : v)
                    std::cout << '\t' << x;
                std::cout << '\n';
            };
            line("SHI", hi);
            sampler.summary_data_int<vsmch::RowMajor>(di.begin());
            line("SDIR", di);
            sampler.summary_data_int<vsmch::ColMajor>(di.begin());
            line("SDIC", di);
            line("SH", hf);
            sampler.summary_data<vsmch::RowMajor>(df.begin());
            line("SDR", df);
            sampler.summary_data<vsmch::ColMajor>(df.begin());
            line("SDC", df);
            std::cout << "STATE\n" << sampler.particle().value();
        }
    } catch (const std::exception &e) {
        std::cerr << "pf_hip: " << e.what() << '\n';
        return 1;
    }
    return 0;
}
