// mcf4d_abi.cpp -- the MCF verification program's C ABI (include/mcf4d.h) on
// top of the HIP engine.  Globals and entry points keep the reference's names
// and meaning (mcf_4D.c:11-471, mcf_4D.h:8-24).
#include "mcf4d.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>

#include "host_mem.h"
#include "host_track.h"
#include "seg4d_engine.h"

#define API extern "C" __attribute__((visibility("default")))

// ---- reference globals (mcf_4D.c:11-34) -------------------------------------
API int kMax = 0, iMax = 0, jMax = 0, lMax = 0;
API double *u = nullptr, *prev_u = nullptr, *exact_u = nullptr;
API double *e = nullptr, *w = nullptr, *n = nullptr, *s = nullptr;
API double *t = nullptr, *b = nullptr, *fw = nullptr, *bw = nullptr;
API int maxNoGSIteration = 0;
API double eps2 = 0;
API int numberOfTimeStep = 0, maxNoOfTimeSteps = 0;
API double mcfTol = 0;
API double tau = 0, h = 0, omega_c = 0, gsTol = 0;

namespace {

s4d::Engine* g_eng = nullptr;
int g_resident = 0;
int g_verbose = 1;
double g_rel = 0.0;
s4d::SolveStats g_last;

double** coef_ptrs[8] = {&e, &w, &n, &s, &t, &b, &fw, &bw};

long nelem() { return (long)iMax * jMax * kMax * lMax; }

bool fail(const char* what) {
    std::string msg = what;
    if (g_eng && !g_eng->err.empty()) msg += " (" + g_eng->err + ")";
    s4d::set_error(msg);
    std::fprintf(stderr, "mcf4d: %s\n", msg.c_str());
    return false;
}

// Host-write tracking of u, prev_u and exact_u in mirror mode (host_track.h):
// off unless switched on by mcf4dSetWriteTracking(1) or MCF4D_TRACK_WRITES=1
// in the environment at the first _allocateMem_mcf.  _mcfTimeStep,
// _exactSolution_mcf and _l2norm_mcf keep what it knows; every other entry
// point (need_engine) first makes all three arrays dirty.
enum { TR_U = 0, TR_PREV = 1, TR_EXACT = 2 };
int g_track = -1;  // -1: not yet read from the environment

int track_slot(int id) {
    if (id == s4d::A_U) return TR_U;
    if (id == s4d::A_PREV) return TR_PREV;
    if (id == s4d::A_EXACT) return TR_EXACT;
    return -1;
}

void track_register() {
    s4d::track::forget_all();
    if (g_track != 1 || !u) return;
    if (!s4d::track::install()) return;
    const size_t bytes = (size_t)nelem() * sizeof(double);
    s4d::track::set(TR_U, u, bytes);
    s4d::track::set(TR_PREV, prev_u, bytes);
    s4d::track::set(TR_EXACT, exact_u, bytes);
}

bool tracked_clean(int id) {
    const int s = track_slot(id);
    return s >= 0 && !g_resident && s4d::track::clean(s);
}

void mark_clean(int id) {
    const int s = track_slot(id);
    if (s >= 0 && !g_resident) s4d::track::mark_clean(s);
}

bool need_engine_tracked(const char* fn) {
    if (g_resident) s4d::track::touch_all();
    if (g_eng && g_eng->ready()) return true;
    std::string m = std::string(fn) + ": library not allocated (call _allocateMem_mcf first)";
    s4d::set_error(m);
    std::fprintf(stderr, "mcf4d: %s\n", m.c_str());
    return false;
}

bool need_engine(const char* fn) {
    s4d::track::touch_all();
    return need_engine_tracked(fn);
}

int resolve(const double* p) {
    if (!p) return -1;
    if (p == u) return s4d::A_U;
    if (p == prev_u) return s4d::A_PREV;
    if (p == exact_u) return s4d::A_EXACT;
    for (int f = 0; f < 8; f++)
        if (p == *coef_ptrs[f]) return -(10 + f);
    return -1;
}

double* device_of(double* hp, s4d::ArrayId spare, bool upload) {
    const int id = resolve(hp);
    if (id >= 0) {
        if (upload && !g_resident && !tracked_clean(id)) {
            if (g_eng->upload((s4d::ArrayId)id, hp)) return nullptr;
            mark_clean(id);
        }
        return g_eng->full((s4d::ArrayId)id);
    }
    if (g_resident || id <= -10) {
        s4d::set_error("pointer is not a library array usable here");
        return nullptr;
    }
    double* d = g_eng->scratch(spare);
    if (!d) return nullptr;
    if (upload && g_eng->upload(spare, hp)) return nullptr;
    return d;
}

s4d::StepParams step_params() {
    s4d::StepParams P;
    P.h = h;
    P.tau = tau;
    P.eps2 = eps2;
    P.omega = omega_c;
    P.gsTol = gsTol;
    P.rel2 = g_rel * g_rel;
    P.cap = maxNoGSIteration;
    return P;
}

// u, prev_u and exact_u cross PCIe at every step in mirror mode: pinned
// (page-locked); the coefficient arrays only on request: calloc.
s4d::HostArrays g_host;

void free_host() {
    s4d::track::forget_all();
    double** all[] = {&u, &prev_u, &exact_u, &e, &w, &n, &s, &t, &b, &fw, &bw};
    for (double** p : all) {
        g_host.free(*p);
        *p = nullptr;
    }
}

}  // namespace

// ---- mcf_4D.c:36-40 ----------------------------------------------------------
API int ijkl(int i, int j, int k, int l) {
    return l * iMax * jMax * kMax + k * iMax * jMax + j * iMax + i;
}

// ---- mcf_4D.c:42-79 ----------------------------------------------------------
API bool _allocateMem_mcf(void) {
    if (iMax < 3 || jMax < 3 || kMax < 3 || lMax < 3) {
        s4d::set_error("_allocateMem_mcf: padded dims must be >= 3");
        return false;
    }
    const long N = nelem();
    // before any allocation: int indices (mcf_4D.c:36-40) reach 2^31 - 1
    if (N >= (1L << 31)) {
        s4d::set_error("_allocateMem_mcf: padded volume exceeds 2^31 elements (32-bit index, as "
                       "the reference)");
        return false;
    }
    free_host();
    double** all[] = {&u, &prev_u, &exact_u, &e, &w, &n, &s, &t, &b, &fw, &bw};
    int q = 0;
    for (double** p : all) {
        // zero-filled like the reference's (mcf_4D.c:42-63 mallocs; its
        // main fills every array before reading it)
        *p = g_host.alloc((size_t)N, q++ < 3);
        if (!*p) {
            free_host();
            return fail("_allocateMem_mcf: host allocation failed");
        }
    }
    if (!g_eng) g_eng = new s4d::Engine();
    if (g_eng->init(iMax, jMax, kMax, lMax, false, true))
        return fail("_allocateMem_mcf: device allocation failed");
    if (g_track < 0) {
        const char* ev = std::getenv("MCF4D_TRACK_WRITES");
        g_track = (ev && ev[0] == '1') ? 1 : 0;
    }
    track_register();
    return true;
}

API bool _deallocateMem_mcf(void) {
    free_host();
    if (g_eng) {
        g_eng->release();
        delete g_eng;
        g_eng = nullptr;
    }
    return true;
}

// ---- mcf_4D.c:81-100 ---------------------------------------------------------
API double _l2norm_mcf(double* a, double* bb) {
    if (!a || !bb) return false;
    if (!need_engine_tracked("_l2norm_mcf")) return 0.0;
    double* da = device_of(a, s4d::A_TMP, true);
    double* db = device_of(bb, s4d::A_TMP2, true);
    if (!da || !db) {
        fail("_l2norm_mcf: cannot resolve arguments");
        return 0.0;
    }
    double r = 0.0;
    if (g_eng->l2norm(da, db, h, &r)) {
        fail("_l2norm_mcf");
        return 0.0;
    }
    return r;
}

// ---- mcf_4D.c:102-205 --------------------------------------------------------
API bool _mcfTimeStep(void) {
    if (!need_engine_tracked("_mcfTimeStep")) return false;
    const double time = tau * numberOfTimeStep;
    // mirror mode: u up (with write tracking only when the host wrote it);
    // the reference's prev_u <- u (mcf_4D.c:113) on host threads while the
    // device computes; u down
    if (!g_resident && !tracked_clean(s4d::A_U) && g_eng->upload(s4d::A_U, u)) {
        s4d::track::touch_all();
        return fail("_mcfTimeStep: upload");
    }
    std::thread prev_copy;
    if (!g_resident) {
        s4d::track::touch(TR_PREV);  // written by the copy
        prev_copy = std::thread([] { s4d::host_copy(prev_u, u, (size_t)nelem()); });
    }
    s4d::SolveStats st;
    const int rc = g_eng->mcf_step(step_params(), time, &st);
    if (prev_copy.joinable()) prev_copy.join();
    // a failed step leaves u as it was; in mirror mode prev_u then holds u^n,
    // as after the reference's first statement
    if (rc) {
        s4d::track::touch_all();
        return fail("_mcfTimeStep");
    }
    g_last = st;
    if (!g_resident) {
        if (g_eng->download(s4d::A_U, u)) {
            s4d::track::touch_all();
            return fail("_mcfTimeStep: download");
        }
        // host and device hold u^(n+1) and u^n alike
        mark_clean(s4d::A_U);
        mark_clean(s4d::A_PREV);
    }
    if (g_verbose) {
        std::printf("Residuum is %e\n", st.R);
        std::printf("Step is %d\n", numberOfTimeStep);
    }
    if (!std::isfinite(st.R)) return fail("_mcfTimeStep: non-finite SOR residual");
    return true;
}

// ---- mcf_4D.c:208-342 --------------------------------------------------------
API bool _gs_coefficients_mcf(void) {
    if (!need_engine("_gs_coefficients_mcf")) return false;
    if (!g_resident && g_eng->upload(s4d::A_U, u)) return fail("_gs_coefficients_mcf: upload");
    if (g_eng->setup(step_params(), false, false)) return fail("_gs_coefficients_mcf");
    if (!g_resident)
        for (int f = 0; f < 8; f++)
            if (g_eng->download_coef(f, *coef_ptrs[f]))
                return fail("_gs_coefficients_mcf: download");
    return g_eng->sync() == 0 ? true : fail("_gs_coefficients_mcf");
}

// ---- mcf_4D.c:344-443 --------------------------------------------------------
API bool _set4DBoundaryToExactValues_mcf(double tt) {
    if (!need_engine("_set4DBoundaryToExactValues_mcf")) return false;
    if (!g_resident && g_eng->upload(s4d::A_U, u)) return fail("_set4DBoundary: upload");
    if (g_eng->mcf_exact(g_eng->full(s4d::A_U), h, tt, true)) return fail("_set4DBoundary");
    if (!g_resident && g_eng->download(s4d::A_U, u)) return fail("_set4DBoundary: download");
    return g_eng->sync() == 0 ? true : fail("_set4DBoundaryToExactValues_mcf");
}

// ---- mcf_4D.c:444-462 --------------------------------------------------------
API void _exactSolution_mcf(double tt) {
    if (!need_engine_tracked("_exactSolution_mcf")) return;
    s4d::track::touch(TR_EXACT);  // its device copy changes here
    if (g_eng->mcf_exact(g_eng->full(s4d::A_EXACT), h, tt, false)) {
        fail("_exactSolution_mcf");
        return;
    }
    if (!g_resident) {
        if (g_eng->download(s4d::A_EXACT, exact_u)) {
            fail("_exactSolution_mcf: download");
            return;
        }
        mark_clean(s4d::A_EXACT);
    }
}

// ---- mcf_4D.c:463-471 --------------------------------------------------------
API void _copy4DdataFromSrcToDest_mcf(double* src, double* dst) {
    if (!src || !dst) return;
    if (!g_resident) {
        std::memcpy(dst, src, (size_t)nelem() * sizeof(double));
        return;
    }
    if (!need_engine("_copy4DdataFromSrcToDest_mcf")) return;
    const int ids = resolve(src), idd = resolve(dst);
    if (ids < 0 || idd < 0 || g_eng->copy((s4d::ArrayId)ids, (s4d::ArrayId)idd) || g_eng->sync())
        fail("_copy4DdataFromSrcToDest_mcf: arrays must be library arrays in resident mode");
}

// ---- extensions ------------------------------------------------------------------
API int mcf4dSetResident(int on) {
    s4d::track::touch_all();
    const int old = g_resident;
    g_resident = on ? 1 : 0;
    return old;
}
API void mcf4dSetRelativeTolerance(double rel) { g_rel = rel > 0.0 ? rel : 0.0; }
API bool mcf4dUpload(double* hp) {
    if (!need_engine("mcf4dUpload")) return false;
    const int id = resolve(hp);
    if (id < 0) return fail("mcf4dUpload: not an uploadable library array");
    return g_eng->upload((s4d::ArrayId)id, hp) == 0 ? true : fail("mcf4dUpload");
}
API bool mcf4dDownload(double* hp) {
    if (!need_engine("mcf4dDownload")) return false;
    const int id = resolve(hp);
    if (id <= -10) return g_eng->download_coef(-(id + 10), hp) == 0 ? true : fail("mcf4dDownload");
    if (id < 0) return fail("mcf4dDownload: not a library array");
    return g_eng->download((s4d::ArrayId)id, hp) == 0 ? true : fail("mcf4dDownload");
}
API int mcf4dSetWriteTracking(int on) {
    const int old = g_track == 1 ? 1 : 0;
    g_track = on ? 1 : 0;
    track_register();
    return old;
}
API int mcf4dWriteTrackingState(void) {
    if (g_track != 1 || !s4d::track::current()) return 0;
    return 1 | (s4d::track::clean(TR_U) ? 2 : 0) | (s4d::track::clean(TR_PREV) ? 4 : 0) |
           (s4d::track::clean(TR_EXACT) ? 8 : 0);
}
API void mcf4dSetVerbose(int on) { g_verbose = on; }
API const char* mcf4dLastError(void) { return s4d::last_error(); }
API int mcf4dLastSweeps(void) { return g_last.z; }
API double mcf4dLastResidual(void) { return g_last.R; }
API bool mcf4dSynchronize(void) { return !g_eng || g_eng->sync() == 0; }
