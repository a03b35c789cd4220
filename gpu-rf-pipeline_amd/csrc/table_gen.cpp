// Build-time generator of the library's constant tables (layouts in tables.hpp). Writes C++
// initialisers of device arrays to stdout, as the exact bit patterns of the float32 / FP16
// values, so the compiler stores them unchanged. Host code; compiled with a clang that knows
// _Float16 (the Makefile uses ROCm's).
//
//   table_gen > tables_data.inc

#include "tables.hpp"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

namespace {

std::uint32_t bits(float v)
{
    std::uint32_t u = 0;
    std::memcpy(&u, &v, sizeof u);
    return u;
}

std::uint16_t bits(_Float16 v)
{
    std::uint16_t u = 0;
    std::memcpy(&u, &v, sizeof u);
    return u;
}

template <class Word>
void emit(const char* type, const char* name, const std::vector<Word>& words)
{
    // External linkage and `used`: no kernel names the arrays (they are reached through the
    // addresses the runtime looks up), so an internal or unused one would be dropped from the
    // device code object, and the runtime could not register its host-side name.
    std::printf("__device__ __attribute__((used)) __align__(16) %s %s[%zu] = {", type, name,
                words.size());
    for (std::size_t i = 0; i < words.size(); ++i) {
        std::printf("%s0x%llxu", i == 0 ? "\n    " : (i % 8 == 0 ? ",\n    " : ", "),
                    static_cast<unsigned long long>(words[i]));
    }
    std::printf("\n};\n\n");
}

// A complex table as interleaved (re, im) float32 words.
template <class Fill>
std::vector<std::uint32_t> complex_table(long long size, Fill fill)
{
    std::vector<std::uint32_t> w(2 * size, 0u);
    fill([&](long long m, double c, double s) {
        w[2 * m] = bits(static_cast<float>(c));
        w[2 * m + 1] = bits(static_cast<float>(s));
    });
    return w;
}

}  // namespace

int main()
{
    std::printf("// Generated at build time by table_gen.cpp from tables.hpp; not a source file.\n\n");
    emit("unsigned int", "rf_table_tw4096",
         complex_table(rf::kTwiddleSize, [](auto put) { rf::tables::twiddle4096(put); }));
    emit("unsigned int", "rf_table_stage4096",
         complex_table(rf::kStage4096Size, [](auto put) { rf::tables::stage4096(put); }));
    emit("unsigned int", "rf_table_16384",
         complex_table(rf::kTw16384Size, [](auto put) { rf::tables::table16384(put); }));
    // hi = v rounded to float32 then FP16; lo = the remainder, rounded the same way
    std::vector<std::uint16_t> dft(2 * rf::kDftHalves);
    for (int i = 0; i < rf::kDftHalves; ++i) {
        const double v = rf::tables::dft64_entry(i);
        const _Float16 hi = static_cast<_Float16>(static_cast<float>(v));
        const _Float16 lo = static_cast<_Float16>(static_cast<float>(v - static_cast<double>(hi)));
        dft[i] = bits(hi);
        dft[rf::kDftHalves + i] = bits(lo);
    }
    emit("unsigned short", "rf_table_dft64", dft);
    return 0;
}
