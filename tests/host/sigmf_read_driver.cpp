// Test driver for the sample-file readers (host/io.cpp), built by tests/test_host_sanitizers.py
// with AddressSanitizer and UndefinedBehaviorSanitizer: reads every path given, prints the
// sample count and rate, or the error the reader throws.
#include "io.hpp"

#include <cstdio>
#include <exception>

int main(int argc, char** argv)
{
    for (int i = 1; i < argc; ++i) {
        try {
            const std::vector<cf32> v = read_cf32(argv[i]);
            std::printf("%zu %g\n", v.size(), sample_rate(argv[i]));
        } catch (const std::exception& e) {
            std::printf("error: %s\n", e.what());
        }
    }
    return 0;
}
