// Internal helpers shared by the host set-up code and the device context.
#pragma once

#include <cstdarg>
#include <cstdio>
#include <string>

#include "morphoshell_amd.h"

namespace ms {

// Thread-local message of the last failed context-free call.
void set_last_error(const std::string& msg);
const std::string& last_error();

inline std::string fmt(const char* f, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, f);
    std::vsnprintf(buf, sizeof(buf), f, ap);
    va_end(ap);
    return std::string(buf);
}

// Compensated (Kahan) sum in the order given, as the reference's kahan_sum
// (kahan_sum.cpp:32-49).
template <class It>
double kahan_sum(It first, It last) {
    volatile double sum = 0.0, c = 0.0;
    for (; first != last; ++first) {
        double y = *first - c;
        double t = sum + y;
        volatile double d = t - sum;
        c = d - y;
        sum = t;
    }
    return sum;
}

}  // namespace ms
