// Prints (as raw little-endian doubles) the first argv[1] draws of
// std::uniform_real_distribution<double>(-1, 1) fed by std::mt19937 seeded with 1:
// the jitter stream of boxTriCollision.cpp:649-659 and 1079-1090.
#include <cstdio>
#include <cstdlib>
#include <random>

int main(int argc, char **argv)
{
    int n = argc > 1 ? std::atoi(argv[1]) : 3000;
    std::mt19937 gen;
    std::uniform_real_distribution<> dis(-1.0, 1.0);
    gen.seed(1);
    for (int i = 0; i < n; ++i) {
        double d = dis(gen);
        std::fwrite(&d, sizeof(double), 1, stdout);
    }
    return 0;
}
