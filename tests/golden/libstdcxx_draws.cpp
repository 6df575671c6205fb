// Writes libstdc++'s own sampling outputs for the weights read from stdin
// (one hex float per line): K indices drawn by
// std::discrete_distribution<std::size_t>(mt19937(seed)), the first four
// std::generate_canonical<double, 53>(mt19937(seed)) uniforms, and
// discrete_distribution::probabilities() (the normalised table), all as hex
// floats / integers.  Run by make_libstdcxx_draws.py.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    const unsigned seed = static_cast<unsigned>(std::strtoul(argv[1], nullptr, 10));
    const int K = std::atoi(argv[2]);
    std::vector<double> p;
    double x;
    while (std::scanf("%la", &x) == 1) p.push_back(x);
    std::mt19937 gen(seed);
    std::discrete_distribution<std::size_t> dd(p.begin(), p.end());
    for (int i = 0; i < K; ++i) std::printf("%zu\n", dd(gen));
    std::mt19937 g2(seed);
    for (int i = 0; i < 4; ++i) std::printf("%a\n", std::generate_canonical<double, 53>(g2));
    for (double q : dd.probabilities()) std::printf("%a\n", q);
    return 0;
}
