// Prints words of PyTorch's Philox-4x32-10 engine (ATen/core/PhiloxRNGEngine.h,
// header only) for the streams named on the command line, one line per
// stream: `seed row nwords` triples in, hex words out.  Used by
// tests/test_philox.py as an independent implementation of the generator.
#include <ATen/core/PhiloxRNGEngine.h>

#include <cstdio>
#include <cstdlib>

int main(int argc, char **argv) {
    for (int i = 1; i + 2 < argc; i += 3) {
        const unsigned long long seed = std::strtoull(argv[i], nullptr, 0);
        const unsigned long long row = std::strtoull(argv[i + 1], nullptr, 0);
        const int n = std::atoi(argv[i + 2]);
        at::philox_engine e(seed, row, 0);
        for (int k = 0; k < n; ++k) std::printf("%08x%c", e(), k + 1 < n ? ' ' : '\n');
    }
    return 0;
}
