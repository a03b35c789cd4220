// Twiddle and DFT tables. The per-device constant tables are initialised device arrays written
// at build time (table_gen.cpp, tables.hpp): each device gets them when the code object loads,
// and the getters only look up their addresses. fill_twiddles builds the per-plan tables of
// longer transforms on the host, in the same float64 arithmetic.

#include "twiddle.hpp"

#include <mutex>

#include "tables_data.inc"

namespace rf {

void fill_twiddles(float2* host, long long count, long long n, long long step)
{
    tables::twiddles(count, n, step, [host](long long m, double c, double s) {
        host[m] = make_float2(static_cast<float>(c), static_cast<float>(s));
    });
}

namespace {

constexpr int kMaxDevices = 64;

// Device address of one table symbol, looked up once per device. The lookup allocates and
// copies nothing.
class TableAddress {
public:
    explicit TableAddress(const void* symbol) : symbol_(symbol) {}

    const void* get()
    {
        int device = 0;
        RF_HIP_CHECK(hipGetDevice(&device));
        if (device < 0 || device >= kMaxDevices) {
            fail(RF_LIMIT, "constant tables: device index out of range");
        }
        std::call_once(once_[device], [&] {
            void* p = nullptr;
            RF_HIP_CHECK(hipGetSymbolAddress(&p, symbol_));
            addr_[device] = p;
        });
        if (addr_[device] == nullptr) {
            fail(RF_HIP_ERROR, "constant tables: symbol lookup failed on an earlier call");
        }
        return addr_[device];
    }

private:
    const void* symbol_;
    std::once_flag once_[kMaxDevices];
    const void* addr_[kMaxDevices] = {};
};

TableAddress g_tw4096(HIP_SYMBOL(rf_table_tw4096));
TableAddress g_stage4096(HIP_SYMBOL(rf_table_stage4096));
TableAddress g_16384(HIP_SYMBOL(rf_table_16384));
TableAddress g_dft64(HIP_SYMBOL(rf_table_dft64));

}  // namespace

const float2* twiddle_table_4096()
{
    return static_cast<const float2*>(g_tw4096.get());
}

const float2* fft4096_stage_table()
{
    return static_cast<const float2*>(g_stage4096.get());
}

const float2* fft16384_table()
{
    return static_cast<const float2*>(g_16384.get());
}

const _Float16* dft64_table()
{
    return static_cast<const _Float16*>(g_dft64.get());
}

}  // namespace rf
