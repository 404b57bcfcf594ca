// Small host-side helpers shared by the library's C++ translation units.
#pragma once
#include <cstdint>
#include <functional>
#include <string>

namespace adds {
// Worker threads used by host loops (ADDS_HOST_THREADS overrides hardware_concurrency, capped at 64).
unsigned host_threads();
// Runs fn(begin, end) over [0, n) in grains of `grain` items on host_threads() threads.
void parallel_for(uint64_t n, uint64_t grain, const std::function<void(uint64_t, uint64_t)>& fn);
// Sets the text adds_last_error_detail() returns on this thread.
void set_error_detail(const std::string& s);
}  // namespace adds
