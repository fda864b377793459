// host_track.h -- host-write tracking of the library's host arrays by page
// protection, for mirror mode.
//
// In mirror mode the host arrays are authoritative, so every entry point
// uploads what it reads.  A driver loop that only calls the library between
// steps (segmentation4D.c:1398-1411: _segTimeStep_, then _l2norm_(prev_u, u))
// would upload u, prev_u and Mmask again although they still equal the device
// copies.  An array whose host copy is known to equal its device copy is
// "clean" and kept read-only; the first CPU write to it faults, and the
// handler makes it writable again and marks it dirty, so it is uploaded the
// next time it is read.  Device copies into the array (DMA) do not fault.
//
// The handler chains to the SIGSEGV disposition in place when it was
// installed, for faults outside the tracked arrays.  An array is protected
// only while this handler is still the installed one; if another handler has
// replaced it since, nothing is protected and every array counts as dirty
// (the plain mirror behaviour).  Writes by the kernel on the process's behalf
// (read(2) straight into a clean array) cannot be caught this way and fail
// with EFAULT: mark the array dirty first (seg4dHostWrote) or switch the
// tracking off.
#pragma once
#include <signal.h>
#include <sys/mman.h>
#include <unistd.h>

#include <atomic>
#include <cstddef>
#include <cstdint>
#include <cstring>

namespace s4d {
namespace track {

constexpr int kSlots = 4;

struct Slot {
    std::atomic<char*> base{nullptr};  // page-aligned start of the array
    std::atomic<size_t> len{0};        // protected length (whole pages)
    std::atomic<int> clean{0};         // 1: read-only, host == device copy
};

inline Slot g_slot[kSlots];
inline struct sigaction g_prev;  // disposition before this handler
inline std::atomic<int> g_installed{0};

inline void on_fault(int sig, siginfo_t* si, void* ctx) {
    char* a = static_cast<char*>(si->si_addr);
    // every slot whose pages hold the address (the last page of one array
    // may be shared with the next allocation)
    bool ours = false;
    for (int s = 0; s < kSlots; s++) {
        char* b = g_slot[s].base.load(std::memory_order_acquire);
        const size_t n = g_slot[s].len.load(std::memory_order_acquire);
        if (b && a >= b && a < b + n) {
            g_slot[s].clean.store(0, std::memory_order_release);
            mprotect(b, n, PROT_READ | PROT_WRITE);
            ours = true;
        }
    }
    if (ours) return;  // the write is retried on a writable page
    // not a tracked array: the disposition in place before
    if (g_prev.sa_flags & SA_SIGINFO) {
        if (g_prev.sa_sigaction) {
            g_prev.sa_sigaction(sig, si, ctx);
            return;
        }
    } else if (g_prev.sa_handler != SIG_DFL && g_prev.sa_handler != SIG_IGN) {
        g_prev.sa_handler(sig);
        return;
    }
    // default action: restored here, taken when the access repeats
    struct sigaction dfl;
    std::memset(&dfl, 0, sizeof dfl);
    dfl.sa_handler = SIG_DFL;
    sigemptyset(&dfl.sa_mask);
    sigaction(sig, &dfl, nullptr);
}

// This handler is the current SIGSEGV disposition.
inline bool current() {
    struct sigaction cur;
    if (sigaction(SIGSEGV, nullptr, &cur) != 0) return false;
    return (cur.sa_flags & SA_SIGINFO) && cur.sa_sigaction == on_fault;
}

// Installs the handler once per library (chaining to the disposition found).
inline bool install() {
    if (g_installed.load()) return current();
    struct sigaction sa;
    std::memset(&sa, 0, sizeof sa);
    sa.sa_sigaction = on_fault;
    sa.sa_flags = SA_SIGINFO;
    sigemptyset(&sa.sa_mask);
    if (sigaction(SIGSEGV, &sa, &g_prev) != 0) return false;
    g_installed.store(1);
    return true;
}

// Makes slot s writable and dirty.
inline void touch(int s) {
    char* b = g_slot[s].base.load();
    const size_t n = g_slot[s].len.load();
    if (b && g_slot[s].clean.exchange(0)) mprotect(b, n, PROT_READ | PROT_WRITE);
}

inline void touch_all() {
    for (int s = 0; s < kSlots; s++) touch(s);
}

// Slot s: [p, p + bytes), dirty.  Only page-aligned arrays of at least 64
// pages are tracked (smaller ones cost little to upload).
inline void set(int s, void* p, size_t bytes) {
    touch(s);
    const long pg = sysconf(_SC_PAGESIZE);
    char* b = static_cast<char*>(p);
    if (!p || pg <= 0 || reinterpret_cast<uintptr_t>(b) % (uintptr_t)pg ||
        bytes < 64 * (size_t)pg) {
        g_slot[s].base.store(nullptr);
        g_slot[s].len.store(0);
        return;
    }
    g_slot[s].len.store((bytes + (size_t)pg - 1) / (size_t)pg * (size_t)pg);
    g_slot[s].base.store(b, std::memory_order_release);
}

inline void forget_all() {
    touch_all();
    for (int s = 0; s < kSlots; s++) {
        g_slot[s].base.store(nullptr);
        g_slot[s].len.store(0);
    }
}

inline bool clean(int s) { return g_slot[s].clean.load(std::memory_order_acquire) != 0; }

// The host copy of slot s now equals its device copy: read-only until the
// next CPU write (only while this handler is installed).
inline void mark_clean(int s) {
    char* b = g_slot[s].base.load();
    const size_t n = g_slot[s].len.load();
    if (!b || !current()) return;
    // clean before the pages turn read-only: a write that faults after the
    // mprotect finds it set and clears it with the protection
    g_slot[s].clean.store(1, std::memory_order_release);
    if (mprotect(b, n, PROT_READ) != 0) g_slot[s].clean.store(0, std::memory_order_release);
}

}  // namespace track
}  // namespace s4d
