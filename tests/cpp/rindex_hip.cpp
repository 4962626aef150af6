// ResampleIndex through the C++ layer (include/vsmch/resample_index.hpp).
// Reads a history from a text file -- one line per iteration, either "I"
// (a step without resampling) or N ancestor indices -- pushes it row by row
// from device arrays, and prints index_matrix<RowMajor>, then
// index_matrix<ColMajor>, then index(i, t) for every i and t, one
// tab-separated line per row.
#include <vsmch/resample_index.hpp>

#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

int main(int argc, char **argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: rindex_hip N history.txt\n");
        return 2;
    }
    const std::size_t N = std::stoul(argv[1]);
    vsmch::ResampleIndex r(N);
    int *d = nullptr;
    vsmch::check_hip(hipMalloc(&d, sizeof(int) * N), "hipMalloc");
    std::ifstream in(argv[2]);
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty())
            continue;
        if (line == "I") {
            r.push_back();
            continue;
        }
        std::istringstream ss(line);
        std::vector<int> row(N);
        for (std::size_t i = 0; i != N; ++i)
            ss >> row[i];
        vsmch::check_hip(hipMemcpy(d, row.data(), sizeof(int) * N, hipMemcpyHostToDevice),
            "hipMemcpy");
        r.push_back(d);
    }
    const std::size_t T = r.iter_size();
    const std::vector<int> rm = r.index_matrix<vsmch::RowMajor>();
    const std::vector<int> cm = r.index_matrix<vsmch::ColMajor>();
    for (std::size_t i = 0; i != N; ++i) {
        for (std::size_t t = 0; t != T; ++t)
            std::printf("%d\t", rm[i * T + t]);
        std::printf("\n");
    }
    for (std::size_t t = 0; t != T; ++t) {
        for (std::size_t i = 0; i != N; ++i)
            std::printf("%d\t", cm[t * N + i]);
        std::printf("\n");
    }
    for (std::size_t i = 0; i != N; ++i) {
        for (std::size_t t = 0; t != T; ++t)
            std::printf("%d\t", r.index(i, t));
        std::printf("\n");
    }
    (void)hipFree(d);
    return 0;
}
