// host_mem.h -- the C ABIs' host arrays: page-locked where they cross PCIe
// every call in mirror mode, zero-filled like the reference's allocations
// (segmentation4D.c:106-133, mcf_4D.c:42-63), and copied on several host
// threads.  One HostArrays per library.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace s4d {

// dst <- src on up to 8 host threads
inline void host_copy(double* dst, const double* src, size_t n) {
    const unsigned nt = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    const size_t per = (n + nt - 1) / nt;
    for (unsigned q = 0; q < nt; q++) {
        const size_t lo = q * per, hi = std::min(n, lo + per);
        if (lo < hi)
            th.emplace_back([=] { std::memcpy(dst + lo, src + lo, (hi - lo) * sizeof(double)); });
    }
    for (auto& x : th) x.join();
}

inline void host_zero(double* p, size_t n) {
    const unsigned nt = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    const size_t per = (n + nt - 1) / nt;
    for (unsigned q = 0; q < nt; q++) {
        const size_t lo = q * per, hi = std::min(n, lo + per);
        if (lo < hi) th.emplace_back([=] { std::memset(p + lo, 0, (hi - lo) * sizeof(double)); });
    }
    for (auto& x : th) x.join();
}

struct HostArrays {
    std::vector<double*> pinned;

    // n zeroed doubles; page-locked when `pin` (pageable if that fails)
    double* alloc(size_t n, bool pin) {
        if (pin) {
            void* p = nullptr;
            if (hipHostMalloc(&p, n * sizeof(double), hipHostMallocDefault) == hipSuccess && p) {
                host_zero(static_cast<double*>(p), n);
                pinned.push_back(static_cast<double*>(p));
                return static_cast<double*>(p);
            }
            (void)hipGetLastError();  // pageable memory instead
        }
        return static_cast<double*>(std::calloc(n, sizeof(double)));
    }

    void free(double* p) {
        if (!p) return;
        auto it = std::find(pinned.begin(), pinned.end(), p);
        if (it != pinned.end()) {
            pinned.erase(it);
            (void)hipHostFree(p);
        } else {
            std::free(p);
        }
    }
};

}  // namespace s4d
