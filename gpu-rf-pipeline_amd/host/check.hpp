#pragma once

// Turns a non-OK rf_status into std::runtime_error carrying the library's message, the
// reference's error convention (src/check.hpp:42-66): main() catches it and exits with 1.
//
//   RF_CHECK(rf_fir(...));

#include "rfhip.h"

#include <stdexcept>
#include <string>

[[noreturn]] inline void throw_rf_error(rf_status status, const char* call, const char* file,
                                        int line)
{
    throw std::runtime_error(std::string(rf_status_string(status)) + ": " + rf_last_error()
                             + "\n  in: " + call + "\n  at: " + file + ":"
                             + std::to_string(line));
}

#define RF_CHECK(call)                                                                        \
    do {                                                                                      \
        const rf_status rf_check_status_ = (call);                                            \
        if (rf_check_status_ != RF_OK) {                                                      \
            throw_rf_error(rf_check_status_, #call, __FILE__, __LINE__);                      \
        }                                                                                     \
    } while (0)
