// Writes random_golden.json: the first draws of libstdc++'s
// std::default_random_engine (default seed) through
// std::uniform_real_distribution<double>(0, 1), the generator and
// distribution the reference's Incident_Wave::set_random_incident_directions
// uses (Incident_Wave.cpp:32-33).  Test fixture only.
//   g++ -O2 -std=c++14 -o /tmp/make_random_golden tests/golden/make_random_golden.cpp
//   /tmp/make_random_golden > tests/golden/random_golden.json
#include <cstdio>
#include <random>

int main()
{
    std::default_random_engine re;
    std::uniform_real_distribution<double> urf{0, 1.0};
    std::printf("{\"engine\": \"std::default_random_engine (default seed)\", "
                "\"distribution\": \"std::uniform_real_distribution<double>(0, 1)\", \"draws\": [");
    for (int i = 0; i < 60; ++i) std::printf("%s%.17g", i ? ", " : "", urf(re));
    std::default_random_engine raw;
    std::printf("], \"engine_outputs\": [");
    for (int i = 0; i < 8; ++i) std::printf("%s%lu", i ? ", " : "", (unsigned long)raw());
    std::printf("]}\n");
    return 0;
}
