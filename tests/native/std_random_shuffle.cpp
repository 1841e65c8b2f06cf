// Drives libstdc++'s own std::random_shuffle and std::mt19937 for the
// random projection order test (tests/test_random_order_pinned.py).
//
// ProjectionSetRandom::reorder (ettention/libReconstruction/src/model/
// projectionset/ProjectionSetRandom.cpp:17-38) fills 0..n-1 and calls
// std::random_shuffle(first, last, fct) with fct(i) = boost::uniform_int<>(0,
// i - 1) on a default-seeded boost::mt19937.  Here the generator functor wraps
// the oracle's engine and uniform_int mapping, so the printed order is what
// libstdc++'s shuffle makes of the oracle's draws.
//
// Output, one line each:
//   engine <1000 outputs of a default-constructed std::mt19937>
//   shuffle <n> <the order for n>          for every n on the command line
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

extern "C" {
uint32_t orc_mt19937_state_words(void);
void orc_mt19937_seed(uint32_t* state, uint32_t seed);
uint32_t orc_boost_uniform_int(uint32_t* state, uint32_t lo, uint32_t hi);
}

namespace {
struct OracleUniform {
    uint32_t* state;
    unsigned operator()(unsigned i) { return orc_boost_uniform_int(state, 0u, i - 1u); }
};
}  // namespace

int main(int argc, char** argv)
{
    std::mt19937 eng;   // default seed 5489, as boost::random::mt19937
    std::printf("engine");
    for (int k = 0; k < 1000; ++k) std::printf(" %u", (unsigned)eng());
    std::printf("\n");
    for (int a = 1; a < argc; ++a) {
        const unsigned n = (unsigned)std::strtoul(argv[a], nullptr, 10);
        std::vector<uint32_t> st(orc_mt19937_state_words());
        orc_mt19937_seed(st.data(), 5489u);
        std::vector<unsigned> idx(n);
        for (unsigned i = 0; i < n; ++i) idx[i] = i;
        OracleUniform fct{st.data()};
        std::random_shuffle(idx.begin(), idx.end(), fct);
        std::printf("shuffle %u", n);
        for (unsigned v : idx) std::printf(" %u", v);
        std::printf("\n");
    }
    return 0;
}
