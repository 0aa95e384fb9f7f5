// ranlux_std_check.cpp -- TEST INFRASTRUCTURE: an independent RANLUX stream
// from libstdc++'s std::ranlux24_base (subtract_with_carry_engine<24, 10, 24>)
// and std::discard_block_engine, used to pin the oracle's RANLUXCL restatement.
// stdin: nskip count pre, then the 24 lagged values x_{i-24} .. x_{i-1} (integers
// in [0, 2^24)) and the carry (0/1).  The engine first discards `pre` values;
// stdout: `count` outputs, one per line.
#include <iostream>
#include <random>
#include <sstream>

int main()
{
    long long nskip, count, pre;
    std::cin >> nskip >> count >> pre;
    std::ostringstream state;
    for (int k = 0; k < 25; k++) {
        long long v;
        std::cin >> v;
        state << v << ' ';
    }
    std::ranlux24_base base;
    std::istringstream(state.str()) >> base;
    base.discard(pre);
    if (nskip == 0) {
        for (long long i = 0; i < count; i++) std::cout << base() << '\n';
        return 0;
    }
    // the block engine's block size is a template parameter: cover the luxury levels used
    auto run = [&](auto eng) {
        for (long long i = 0; i < count; i++) std::cout << eng() << '\n';
    };
    switch (nskip) {
    case 24: run(std::discard_block_engine<std::ranlux24_base, 48, 24>(base)); break;
    case 76: run(std::discard_block_engine<std::ranlux24_base, 100, 24>(base)); break;
    case 200: run(std::discard_block_engine<std::ranlux24_base, 224, 24>(base)); break;
    case 380: run(std::discard_block_engine<std::ranlux24_base, 404, 24>(base)); break;
    default: return 2;
    }
    return 0;
}
