// seg4d_abi.cpp -- the segmentation program's C ABI (include/seg4d.h) on top
// of the HIP engine.  Globals and entry points keep the reference's names
// and meaning (segmentation4D.c:13-60, segmentation4D.h:8-62).
#include "seg4d.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "host_mem.h"
#include "host_track.h"
#include "seg4d_batch.h"
#include "seg4d_engine.h"
#include "seg4d_net.h"

#define API extern "C" __attribute__((visibility("default")))

// ---- reference globals (segmentation4D.c:13-60) ----------------------------
API int kMax = 0, iMax = 0, jMax = 0, lMax = 0;
API double *u = nullptr, *prev_u = nullptr, *i0 = nullptr, *Mmask = nullptr;
API double *e = nullptr, *w = nullptr, *n = nullptr, *s = nullptr;
API double *t = nullptr, *b = nullptr, *fw = nullptr, *bw = nullptr;
API double *g_e = nullptr, *g_w = nullptr, *g_n = nullptr, *g_s = nullptr;
API double *g_t = nullptr, *g_b = nullptr, *g_fw = nullptr, *g_bw = nullptr;
API int* no_of_centers = nullptr;
API double **center_x = nullptr, **center_y = nullptr, **center_z = nullptr;
API double spacing_x = 0, spacing_y = 0, spacing_z = 0;
API double origin_x = 0, origin_y = 0, origin_z = 0;
API int maxNoGSIteration = 0;
API double eps2 = 0, K = 0;
API int numberOfTimeStep = 0, maxNoOfTimeSteps = 0;
API double segTol = 0;
API double tau = 0, h = 0, omega_c = 0, gsTol = 0, omega_sigma = 0;
API double theta_img = 0, rho_thr = 0, alpha = 0, beta = 0;
API int mod = 0;
API double sigma = 0;
API double maxNoGSIteration_sigma = 0;
API int timeStepsNum = 0;
API double ballRadius = 0;

namespace {

s4d::Engine* g_eng = nullptr;
s4d::Net* g_net = nullptr;   // learned warm start (survives re-allocation)
s4d::Batch* g_batch = nullptr;  // batched form (independent of the globals' arrays)
int g_resident = 0;
int g_verbose = 1;
double g_rel = 0.0;
bool g_g_dirty = true;       // host g_* not yet on the device
int g_centre_frames = 0;     // frames the centre arrays were allocated for
s4d::SolveStats g_last;

double** coef_ptrs[8] = {&e, &w, &n, &s, &t, &b, &fw, &bw};
double** g_ptrs[8] = {&g_e, &g_w, &g_n, &g_s, &g_t, &g_b, &g_fw, &g_bw};

long nelem() { return (long)iMax * jMax * kMax * lMax; }

bool fail(const char* what) {
    std::string msg = what;
    if (g_eng && !g_eng->err.empty()) msg += " (" + g_eng->err + ")";
    s4d::set_error(msg);
    std::fprintf(stderr, "seg4d: %s\n", msg.c_str());
    return false;
}

// Host-write tracking of u, prev_u and Mmask in mirror mode (host_track.h):
// off unless switched on by seg4dSetWriteTracking(1) or SEG4D_TRACK_WRITES=1
// in the environment at the first _allocateMem_.
enum { TR_U = 0, TR_PREV = 1, TR_MASK = 2 };
int g_track = -1;  // -1: not yet read from the environment

int track_slot(int id) {
    if (id == s4d::A_U) return TR_U;
    if (id == s4d::A_PREV) return TR_PREV;
    if (id == s4d::A_MASK) return TR_MASK;
    return -1;
}

// Tracked arrays whose host copy is known to equal the device copy skip
// their upload.  Only _segTimeStep_ and _l2norm_ keep that knowledge; every
// other entry point (need_engine) first makes all three arrays dirty, since
// it may change the device copies.
void track_register() {
    s4d::track::forget_all();
    if (g_track != 1 || !u) return;
    if (!s4d::track::install()) return;
    const size_t bytes = (size_t)nelem() * sizeof(double);
    s4d::track::set(TR_U, u, bytes);
    s4d::track::set(TR_PREV, prev_u, bytes);
    s4d::track::set(TR_MASK, Mmask, bytes);
}

bool tracked_clean(int id) {
    const int s = track_slot(id);
    return s >= 0 && !g_resident && s4d::track::clean(s);
}

void mark_clean(int id) {
    const int s = track_slot(id);
    if (s >= 0 && !g_resident) s4d::track::mark_clean(s);
}

// For the tracked entry points: the engine is usable; tracking knowledge kept
// (dropped in resident mode, where the device copies change unseen).
bool need_engine_tracked(const char* fn) {
    if (g_resident) s4d::track::touch_all();
    if (g_eng && g_eng->ready()) return true;
    std::string m = std::string(fn) + ": library not allocated (call _allocateMem_ first)";
    s4d::set_error(m);
    std::fprintf(stderr, "seg4d: %s\n", m.c_str());
    return false;
}

bool need_engine(const char* fn) {
    s4d::track::touch_all();
    return need_engine_tracked(fn);
}

// Library array id of a host pointer, -1 if foreign, -(10+f) for coef f.
int resolve(const double* p) {
    if (!p) return -1;
    if (p == u) return s4d::A_U;
    if (p == prev_u) return s4d::A_PREV;
    if (p == i0) return s4d::A_I0;
    if (p == Mmask) return s4d::A_MASK;
    for (int f = 0; f < 8; f++)
        if (p == *g_ptrs[f]) return s4d::A_G0 + f;
    for (int f = 0; f < 8; f++)
        if (p == *coef_ptrs[f]) return -(10 + f);
    return -1;
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

s4d::Balls collect_balls() {
    s4d::Balls B;
    B.radius = ballRadius;
    B.count.assign(lMax, 0);
    B.offset.assign(lMax, 0);
    for (int l = 0; l < lMax; l++) {
        int c = (no_of_centers && l < g_centre_frames) ? no_of_centers[l] : 0;
        if (c < 0) c = 0;
        B.offset[l] = (int)B.cx.size();
        if (l >= 1 && l <= lMax - 2) {
            B.count[l] = c;
            for (int q = 0; q < c; q++) {
                B.cx.push_back(center_x[l][q]);
                B.cy.push_back(center_y[l][q]);
                B.cz.push_back(center_z[l][q]);
            }
        }
    }
    B.total = (int)B.cx.size();
    return B;
}

// Device array behind a host pointer.  Library arrays map to their mirror;
// in mirror mode a foreign host array is staged in scratch `spare`.
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

bool download_to(double* hp, const double* d) {
    const int id = resolve(hp);
    if (id >= 0) return g_eng->download((s4d::ArrayId)id, hp) == 0;
    // staged foreign array: find which scratch holds it
    for (s4d::ArrayId a : {s4d::A_TMP, s4d::A_TMP2})
        if (g_eng->full(a) == d) return g_eng->download(a, hp) == 0;
    return false;
}

bool upload_g_if_dirty() {
    if (g_resident || !g_g_dirty) return true;
    for (int f = 0; f < 8; f++)
        if (g_eng->upload((s4d::ArrayId)(s4d::A_G0 + f), *g_ptrs[f])) return false;
    g_g_dirty = false;
    return true;
}

// Host arrays that cross PCIe in mirror mode -- u, prev_u and Mmask every
// time step, i0 and the g arrays with the edge indicator -- are pinned
// (page-locked), so the copies run as DMA at link speed instead of through
// the runtime's staging buffers; the others, and any array whose pinning
// fails, come from calloc.  Both are zero-filled, as the reference's
// allocation (segmentation4D.c:106-133).
s4d::HostArrays g_host;

void free_host() {
    s4d::track::forget_all();
    double** all[] = {&u, &prev_u, &i0, &Mmask, &e, &w, &n, &s, &t, &b, &fw, &bw,
                      &g_e, &g_w, &g_n, &g_s, &g_t, &g_b, &g_fw, &g_bw};
    for (double** p : all) {
        g_host.free(*p);
        *p = nullptr;
    }
    double*** cs[] = {&center_x, &center_y, &center_z};
    for (double*** c : cs) {
        if (*c) {
            for (int l = 0; l < g_centre_frames; l++) std::free((*c)[l]);
            std::free(*c);
        }
        *c = nullptr;
    }
    g_centre_frames = 0;
}

}  // namespace

// ---- segmentation4D.c:61-70 --------------------------------------------------
API int ijkl(int i, int j, int k, int l) {
    return l * iMax * jMax * kMax + k * iMax * jMax + j * iMax + i;
}

API double _gFunction_(double value, double coef) { return 1.0 / (1 + coef * value); }

// ---- segmentation4D.c:71-192 -------------------------------------------------
API bool _allocateMem_(void) {
    if (iMax < 3 || jMax < 3 || kMax < 3 || lMax < 3) {
        s4d::set_error("_allocateMem_: padded dims must be >= 3");
        return false;
    }
    const long N = nelem();
    // checked before any allocation: the reference's int index (ijkl,
    // segmentation4D.c:61-65) and the device kernels' 32-bit indices cover
    // at most 2^31 - 1 padded elements
    if (N >= (1L << 31)) {
        s4d::set_error("_allocateMem_: padded volume exceeds 2^31 elements (32-bit index, as the "
                       "reference)");
        return false;
    }
    free_host();
    double** all[] = {&u, &prev_u, &i0, &Mmask, &e, &w, &n, &s, &t, &b, &fw, &bw,
                      &g_e, &g_w, &g_n, &g_s, &g_t, &g_b, &g_fw, &g_bw};
    // pinned: u, prev_u, i0, Mmask and the g arrays (the coefficient arrays
    // cross PCIe only on request)
    int q = 0;
    for (double** p : all) {
        const bool pin = q < 4 || q >= 12;
        q++;
        *p = g_host.alloc((size_t)N, pin);
        if (!*p) {
            free_host();
            return fail("_allocateMem_: host allocation failed");
        }
    }
    center_x = (double**)std::calloc((size_t)lMax, sizeof(double*));
    center_y = (double**)std::calloc((size_t)lMax, sizeof(double*));
    center_z = (double**)std::calloc((size_t)lMax, sizeof(double*));
    if (!center_x || !center_y || !center_z) return fail("_allocateMem_: centre arrays");
    g_centre_frames = lMax;
    for (int l = 0; l < lMax; l++) {
        int c = no_of_centers ? no_of_centers[l] : 0;
        if (c < 1) c = 1;
        center_x[l] = (double*)std::calloc((size_t)c, sizeof(double));
        center_y[l] = (double*)std::calloc((size_t)c, sizeof(double));
        center_z[l] = (double*)std::calloc((size_t)c, sizeof(double));
        if (!center_x[l] || !center_y[l] || !center_z[l])
            return fail("_allocateMem_: centre arrays");
    }
    if (!g_eng) g_eng = new s4d::Engine();
    if (g_eng->init(iMax, jMax, kMax, lMax, true, false))
        return fail("_allocateMem_: device allocation failed");
    g_eng->net_ = g_net;
    g_g_dirty = true;
    if (g_track < 0) {
        const char* ev = std::getenv("SEG4D_TRACK_WRITES");
        g_track = (ev && ev[0] == '1') ? 1 : 0;
    }
    track_register();
    return true;
}

API bool _deallocateMem_(void) {
    free_host();
    if (g_eng) {
        g_eng->release();
        delete g_eng;
        g_eng = nullptr;
    }
    return true;
}

// ---- segmentation4D.c:194-213 ------------------------------------------------
API double _l2norm_(double* a, double* bb) {
    if (!a || !bb) return false;
    if (!need_engine_tracked("_l2norm_")) return 0.0;
    double* da = device_of(a, s4d::A_TMP, true);
    double* db = device_of(bb, s4d::A_TMP2, true);
    if (!da || !db) {
        fail("_l2norm_: cannot resolve arguments");
        return 0.0;
    }
    double r = 0.0;
    if (g_eng->l2norm(da, db, h, &r)) {
        fail("_l2norm_");
        return 0.0;
    }
    return r;
}

// ---- segmentation4D.c:446-580 ------------------------------------------------
API bool _gs_coefficients_(void) {
    if (!need_engine("_gs_coefficients_")) return false;
    if (!g_resident && g_eng->upload(s4d::A_U, u)) return fail("_gs_coefficients_: upload");
    if (!upload_g_if_dirty()) return fail("_gs_coefficients_: upload g");
    if (g_eng->setup(step_params(), true, false)) return fail("_gs_coefficients_");
    if (!g_resident)
        for (int f = 0; f < 8; f++)
            if (g_eng->download_coef(f, *coef_ptrs[f])) return fail("_gs_coefficients_: download");
    return g_eng->sync() == 0 ? true : fail("_gs_coefficients_");
}

// ---- segmentation4D.c:215-315 ------------------------------------------------
API bool _segTimeStep_(void) {
    if (!need_engine_tracked("_segTimeStep_")) return false;
    if (!upload_g_if_dirty()) return fail("_segTimeStep_: upload g");
    // Mirror mode: u goes up before the step; Mmask, read only by the step's
    // epilogue, goes up beside it on a second stream; the reference's
    // prev_u <- u (segmentation4D.c:224) is done on the host while the device
    // computes; after the step only u comes down.  With write tracking, u and
    // Mmask go up only when the host wrote them since they last matched the
    // device copies.
    std::thread prev_copy;
    if (!g_resident) {
        if ((!tracked_clean(s4d::A_U) && g_eng->upload(s4d::A_U, u)) ||
            (!tracked_clean(s4d::A_MASK) && g_eng->upload_side(s4d::A_MASK, Mmask))) {
            g_eng->finish_side();
            s4d::track::touch_all();
            return fail("_segTimeStep_: upload");
        }
        s4d::track::touch(TR_PREV);  // written by the copy below
        prev_copy = std::thread([] { s4d::host_copy(prev_u, u, (size_t)nelem()); });
    }
    const s4d::Balls balls = collect_balls();
    s4d::SolveStats st;
    const int rc = g_eng->seg_step(step_params(), &balls, &st);
    if (prev_copy.joinable()) prev_copy.join();
    const int rs = g_eng->finish_side();
    // a failed step leaves u as it was; in mirror mode prev_u then holds u^n,
    // as after the reference's first statement
    if (rc || rs) {
        s4d::track::touch_all();
        return fail("_segTimeStep_");
    }
    g_last = st;
    if (!g_resident) {
        if (g_eng->download(s4d::A_U, u)) {
            s4d::track::touch_all();
            return fail("_segTimeStep_: download");
        }
        // host and device now hold u^(n+1), u^n and the mask alike
        mark_clean(s4d::A_U);
        mark_clean(s4d::A_PREV);
        mark_clean(s4d::A_MASK);
    }
    if (g_verbose) {
        std::printf("Step is %d\n", numberOfTimeStep);
        std::printf("Mean square residue is %e and total number of RBSOR is %d\n", st.R, st.z);
    }
    if (!std::isfinite(st.R)) return fail("_segTimeStep_: non-finite SOR residual");
    return true;
}

// ---- segmentation4D.c:337-409 ------------------------------------------------
static bool presmooth_device(double* d) {
    int z = 0;
    double R = 0.0;
    if (g_eng->presmooth(d, sigma, h, omega_sigma, gsTol, g_rel * g_rel,
                         maxNoGSIteration_sigma, &z, &R))
        return false;
    g_last.z = z;
    g_last.R = R;
    if (g_verbose)
        std::printf("Presmoothing mean square residue is %e and total number of RBSOR is %d\n",
                    R, z);
    return true;
}

API bool _presmoothStep_(double* a) {
    if (!a) return false;
    if (!need_engine("_presmoothStep_")) return false;
    double* d = device_of(a, s4d::A_TMP, true);
    if (!d) return fail("_presmoothStep_: cannot resolve argument");
    if (!presmooth_device(d)) return fail("_presmoothStep_");
    if (!g_resident) {
        if (!download_to(a, d) || g_eng->download(s4d::A_PREV, prev_u))
            return fail("_presmoothStep_: download");
    }
    if (!std::isfinite(g_last.R)) return fail("_presmoothStep_: non-finite residual");
    return true;
}

// ---- segmentation4D.c:582-823 ------------------------------------------------
API bool _gf_inputImage_plus_thresholdedImage_(double* a, double* bb) {
    if (!a || !bb) return false;
    if (!need_engine("_gf_inputImage_plus_thresholdedImage_")) return false;
    double* da = device_of(a, s4d::A_TMP, true);
    double* db = device_of(bb, s4d::A_TMP2, true);
    if (!da || !db || da == db) return fail("_gf_: cannot resolve arguments");
    if (!presmooth_device(da) || !presmooth_device(db)) return fail("_gf_: presmoothing");
    // a finite presmoothing residual of b means every value of b is finite
    if (g_eng->gfun(da, db, h, K, theta_img, rho_thr, std::isfinite(g_last.R)))
        return fail("_gf_: edge indicator");
    if (!g_resident) {
        if (!download_to(a, da) || !download_to(bb, db) ||
            g_eng->download(s4d::A_PREV, prev_u))
            return fail("_gf_: download");
        for (int f = 0; f < 8; f++)
            if (g_eng->download((s4d::ArrayId)(s4d::A_G0 + f), *g_ptrs[f]))
                return fail("_gf_: download g");
    }
    g_g_dirty = false;
    return g_eng->sync() == 0 ? true : fail("_gf_");
}

// ---- segmentation4D.c:826-864 ------------------------------------------------
API bool _reflection4Ddata_(double* a) {
    if (!a) return false;
    if (!need_engine("_reflection4Ddata_")) return false;
    double* d = device_of(a, s4d::A_TMP, true);
    if (!d) return fail("_reflection4Ddata_: cannot resolve argument");
    if (g_eng->reflect(d)) return fail("_reflection4Ddata_");
    if (!g_resident && !download_to(a, d)) return fail("_reflection4Ddata_: download");
    return g_eng->sync() == 0 ? true : fail("_reflection4Ddata_");
}

// ---- segmentation4D.c:866-874 ------------------------------------------------
API void _copy4DdataFromSrcToDest_(double* src, double* dst) {
    if (!src || !dst) return;
    const int ids = resolve(src), idd = resolve(dst);
    if (!g_resident) {
        std::memcpy(dst, src, (size_t)nelem() * sizeof(double));
        if (idd >= s4d::A_G0 && idd <= s4d::A_G7) g_g_dirty = true;
        return;
    }
    if (!need_engine("_copy4DdataFromSrcToDest_")) return;
    if (ids < 0 || idd < 0 || g_eng->copy((s4d::ArrayId)ids, (s4d::ArrayId)idd) || g_eng->sync())
        fail("_copy4DdataFromSrcToDest_: arrays must be library arrays in resident mode");
}

// ---- segmentation4D.c:1020-1104 ----------------------------------------------
API bool _rescale4DdataToZeroOne_(double* a) {
    if (!a) return false;
    if (!need_engine("_rescale4DdataToZeroOne_")) return false;
    double* d = device_of(a, s4d::A_TMP, true);
    if (!d) return fail("_rescale4DdataToZeroOne_: cannot resolve argument");
    if (!g_resident && g_eng->upload(s4d::A_MASK, Mmask)) return fail("_rescale: upload mask");
    if (g_eng->rescale(d, collect_balls(), true)) return fail("_rescale4DdataToZeroOne_");
    if (!g_resident && !download_to(a, d)) return fail("_rescale4DdataToZeroOne_: download");
    return g_eng->sync() == 0 ? true : fail("_rescale4DdataToZeroOne_");
}

// ---- segmentation4D.c:317-335 ------------------------------------------------
API double _mass4D_(double* a) {
    if (!a) return false;
    if (!need_engine("_mass4D_")) return 0.0;
    double* d = device_of(a, s4d::A_TMP, true);
    double r = 0.0;
    if (!d || g_eng->mass(d, h, &r)) {
        fail("_mass4D_");
        return 0.0;
    }
    return r;
}

// ---- segmentation4D.c:876-969 ------------------------------------------------
static void exit_on_null(const void* p, const char* fn) {
    if (p) return;
    std::fprintf(stderr, "seg4d: %s: NULL image\n", fn);
    std::exit(1);
}

API double _get4DMax_(double* a) {
    exit_on_null(a, "_get4DMax_");
    if (!need_engine("_get4DMax_")) return 0.0;
    double* d = device_of(a, s4d::A_TMP, true);
    double r = 0.0;
    if (!d || g_eng->max_interior(d, &r)) {
        fail("_get4DMax_");
        return 0.0;
    }
    return r;
}

static double ball_extreme(double* a, int l, double cx, double cy, double cz, double R,
                           bool want_max, const char* fn) {
    exit_on_null(a, fn);
    if (!need_engine(fn)) return 0.0;
    double* d = device_of(a, s4d::A_TMP, true);
    double mn = 0.0, mx = 0.0;
    if (!d || g_eng->ball_minmax(d, l, cx, cy, cz, R, &mn, &mx)) {
        fail(fn);
        return 0.0;
    }
    return want_max ? mx : mn;
}

API double _get3DMaxWithinR_(double* a, int l, double cx, double cy, double cz, double R) {
    return ball_extreme(a, l, cx, cy, cz, R, true, "_get3DMaxWithinR_");
}

API double _get3DMinWithinR_(double* a, int l, double cx, double cy, double cz, double R) {
    return ball_extreme(a, l, cx, cy, cz, R, false, "_get3DMinWithinR_");
}

// ---- segmentation4D.c:971-1018 -----------------------------------------------
API bool _locallyThreshold4Ddata_(double* in, double* thr) {
    if (!in || !thr) return false;
    if (!need_engine("_locallyThreshold4Ddata_")) return false;
    double* din = device_of(in, s4d::A_TMP, true);
    double* dthr = thr == in ? din : device_of(thr, s4d::A_TMP2, true);
    if (!din || !dthr) return fail("_locallyThreshold4Ddata_: cannot resolve arguments");
    if (g_eng->threshold(din, dthr, collect_balls(), alpha, beta))
        return fail("_locallyThreshold4Ddata_");
    if (!g_resident && !download_to(thr, dthr)) return fail("_locallyThreshold4Ddata_: download");
    return g_eng->sync() == 0 ? true : fail("_locallyThreshold4Ddata_");
}

// ---- segmentation4D.c:1106-1148 ----------------------------------------------
API bool _generate4DsegmFunct_(double* a) {
    if (!a) return false;
    if (!need_engine("_generate4DsegmFunct_")) return false;
    double* d = device_of(a, s4d::A_TMP, true);
    if (!d) return fail("_generate4DsegmFunct_: cannot resolve argument");
    if (!g_resident && a != Mmask && g_eng->upload(s4d::A_MASK, Mmask))
        return fail("_generate4DsegmFunct_: upload mask");
    if (g_eng->seed(d, g_eng->full(s4d::A_MASK), collect_balls()))
        return fail("_generate4DsegmFunct_");
    if (!g_resident) {
        if (!download_to(a, d) || g_eng->download(s4d::A_MASK, Mmask))
            return fail("_generate4DsegmFunct_: download");
    }
    return g_eng->sync() == 0 ? true : fail("_generate4DsegmFunct_");
}

// ---- extensions ------------------------------------------------------------------
API int seg4dSetResident(int on) {
    s4d::track::touch_all();
    const int old = g_resident;
    g_resident = on ? 1 : 0;
    return old;
}

API void seg4dSetRelativeTolerance(double rel) { g_rel = rel > 0.0 ? rel : 0.0; }

API bool seg4dUpload(double* hp) {
    if (!need_engine("seg4dUpload")) return false;
    const int id = resolve(hp);
    if (id < 0) return fail("seg4dUpload: not an uploadable library array");
    if (g_eng->upload((s4d::ArrayId)id, hp)) return fail("seg4dUpload");
    if (id >= s4d::A_G0 && id <= s4d::A_G7) {
        // the g set is clean once all eight have been uploaded explicitly
        g_g_dirty = false;
    }
    return true;
}

API bool seg4dDownload(double* hp) {
    if (!need_engine("seg4dDownload")) return false;
    const int id = resolve(hp);
    if (id <= -10) return g_eng->download_coef(-(id + 10), hp) == 0 ? true : fail("seg4dDownload");
    if (id < 0) return fail("seg4dDownload: not a library array");
    return g_eng->download((s4d::ArrayId)id, hp) == 0 ? true : fail("seg4dDownload");
}

API void seg4dHostWrote(double* hp) {
    const int id = resolve(hp);
    const int slot = track_slot(id);
    if (slot >= 0) s4d::track::touch(slot);
    if (id >= s4d::A_G0 && id <= s4d::A_G7) g_g_dirty = true;
}

API int seg4dSetWriteTracking(int on) {
    const int old = g_track == 1 ? 1 : 0;
    g_track = on ? 1 : 0;
    track_register();
    return old;
}

API int seg4dWriteTrackingState(void) {
    if (g_track != 1 || !s4d::track::current()) return 0;
    return 1 | (s4d::track::clean(TR_U) ? 2 : 0) | (s4d::track::clean(TR_PREV) ? 4 : 0) |
           (s4d::track::clean(TR_MASK) ? 8 : 0);
}

API void seg4dSetVerbose(int on) { g_verbose = on; }
API const char* seg4dLastError(void) { return s4d::last_error(); }
API int seg4dLastSweeps(void) { return g_last.z; }
API double seg4dLastResidual(void) { return g_last.R; }
API double seg4dLastTolerance(void) { return g_last.tol; }

API void seg4dProfile(int on) {
    if (g_eng) g_eng->prof.on = on != 0;
}

API void seg4dProfileReset(void) {
    if (g_eng) {
        const bool on = g_eng->prof.on;
        g_eng->prof = s4d::Profile();
        g_eng->prof.on = on;
    }
}

API int seg4dProfileRead(double* out, int nout) {
    if (!g_eng || !out) return 0;
    const s4d::Profile& p = g_eng->prof;
    const double v[25] = {p.sweep_ms[0], (double)p.sweep_launches[0],
                          p.sweep_ms[1], (double)p.sweep_launches[1],
                          p.sweep_ms[2], (double)p.sweep_launches[2],
                          p.setup_ms, (double)p.setup_launches, p.step_ms, (double)p.steps,
                          (double)p.voxel_updates, (double)p.iterations,
                          (double)p.nred, (double)p.nblack,
                          p.sweep_ms[3], (double)p.sweep_launches[3],
                          p.sweep_ms[4], (double)p.sweep_launches[4],
                          p.sweep_ms[5], (double)p.sweep_launches[5],
                          p.defect_ms, (double)p.defect_launches, (double)p.rounds,
                          p.copy_ms, (double)p.copy_launches};
    int k = nout < 25 ? nout : 25;
    for (int q = 0; q < k; q++) out[q] = v[q];
    return k;
}

API int seg4dSetPrecision(int mode) {
    if (!g_eng) return -1;
    const int old = g_eng->precision_;
    g_eng->precision_ = (mode == 1) ? 1 : 0;
    return old;
}

API void seg4dSetRefinement(int max_rounds, double inner_floor) {
    if (!g_eng) return;
    if (max_rounds > 0) g_eng->max_rounds_ = max_rounds;
    if (inner_floor > 0) g_eng->inner_floor2_ = inner_floor * inner_floor;
}

API void seg4dSetInnerFraction(double frac) {
    if (g_eng && frac > 0.0 && frac < 1.0) g_eng->inner_frac_ = frac;
}

API double seg4dSetInnerOmega(double omega) {
    if (!g_eng) return -1.0;
    const double old = g_eng->inner_omega_;
    if (omega == 0.0 || (omega > 0.0 && omega < 2.0)) g_eng->inner_omega_ = omega;
    return old;
}

API int seg4dLastRounds(void) { return g_last.rounds; }

API int seg4dSetWarmStart(int mode) {
    if (!g_eng) return -1;
    if (mode == 2 && !(g_net && g_net->ready())) {
        s4d::set_error("seg4dSetWarmStart: no network (seg4dNetInit or seg4dNetLoad first)");
        return -1;
    }
    const int old = g_eng->warm_;
    g_eng->warm_ = (mode >= 1 && mode <= 4) ? mode : 0;
    return old;
}

// ---- learned warm start ------------------------------------------------------------
static bool need_net(const char* fn) {
    if (g_net && g_net->ready()) return true;
    s4d::set_error(std::string(fn) + ": no network (seg4dNetInit or seg4dNetLoad first)");
    return false;
}

static int net_fail(const char* fn) {
    std::string m = std::string(fn) + ": " + (g_net ? g_net->err : std::string("no network"));
    s4d::set_error(m);
    std::fprintf(stderr, "seg4d: %s\n", m.c_str());
    return -1;
}

API int seg4dNetInit(uint64_t seed) {
    if (!g_net) g_net = new s4d::Net();
    if (g_net->init(seed)) return net_fail("seg4dNetInit");
    if (g_eng) g_eng->net_ = g_net;
    return 0;
}

API int seg4dNetInitShape(uint64_t seed, int layers, int base) {
    if (!g_net) g_net = new s4d::Net();
    if (g_net->init(seed, layers, base)) return net_fail("seg4dNetInitShape");
    if (g_eng) g_eng->net_ = g_net;
    return 0;
}

API int seg4dNetShape(int* layers, int* base) {
    if (!need_net("seg4dNetShape")) return -1;
    if (layers) *layers = g_net->layers();
    if (base) *base = g_net->base();
    return 0;
}

API int seg4dNetLoad(const char* path) {
    if (!g_net) g_net = new s4d::Net();
    if (g_net->load(path)) return net_fail("seg4dNetLoad");
    if (g_eng) g_eng->net_ = g_net;
    return 0;
}

API int seg4dNetSave(const char* path) {
    if (!need_net("seg4dNetSave")) return -1;
    return g_net->save(path) ? net_fail("seg4dNetSave") : 0;
}

API int seg4dNetRecord(int new_run) {
    if (!need_net("seg4dNetRecord") || !need_engine("seg4dNetRecord")) return -1;
    if (!g_resident && g_eng->upload(s4d::A_U, u)) {
        fail("seg4dNetRecord: upload");
        return -1;
    }
    return g_net->record(*g_eng, new_run != 0) ? net_fail("seg4dNetRecord") : 0;
}

API long seg4dNetPoolPairs(void) { return g_net ? g_net->pool_pairs() : 0; }

API double seg4dNetTrain(int iters, int batch, double lr, uint64_t seed, int log_every) {
    if (!need_net("seg4dNetTrain")) return -1.0;
    double loss = 0.0;
    if (g_net->train(iters, batch, (float)lr, seed, log_every, &loss)) {
        net_fail("seg4dNetTrain");
        return -1.0;
    }
    return loss;
}

API int seg4dNetParams(float* out, int n) {
    if (!need_net("seg4dNetParams")) return -1;
    if (!out || n <= 0) return g_net->nparam;
    const int m = n < g_net->nparam ? n : g_net->nparam;
    if (hipMemcpy(out, g_net->wm_, (size_t)m * 4, hipMemcpyDeviceToHost) != hipSuccess) return -1;
    return g_net->nparam;
}

API int seg4dNetSetParams(const float* in, int n, float s_d, float s_t) {
    if (!need_net("seg4dNetSetParams")) return -1;
    if (!in || n != g_net->nparam) {
        s4d::set_error("seg4dNetSetParams: need exactly the network's parameter count");
        return -1;
    }
    // write the FP32 master weights, then re-derive the BF16 and transposed
    // copies the kernels read
    if (hipMemcpy(g_net->wm_, in, (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess)
        return -1;
    g_net->s_d = s_d;
    g_net->s_t = s_t;
    return g_net->refresh() ? net_fail("seg4dNetSetParams") : 0;
}

API int seg4dNetGrad(int I, int J, int K, int L, int P, const uint16_t* feat_bf16,
                     const float* target, const unsigned char* lmask, float* grad_out,
                     double* loss) {
    if (!need_net("seg4dNetGrad")) return -1;
    const size_t pts = (size_t)I * J * K * L;
    void* df = nullptr;
    float* dt = nullptr;
    unsigned char* dm = nullptr;
    int rc = 0;
    if (hipMalloc(&df, pts * 16) != hipSuccess || hipMalloc((void**)&dt, pts * 4) != hipSuccess ||
        hipMalloc((void**)&dm, pts) != hipSuccess ||
        hipMemcpy(df, feat_bf16, pts * 16, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dt, target, pts * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dm, lmask, pts, hipMemcpyHostToDevice) != hipSuccess) {
        s4d::set_error("seg4dNetGrad: device buffers");
        rc = -1;
    }
    if (!rc && g_net->grad_batch(I, J, K, L, P, df, dt, dm, loss)) rc = net_fail("seg4dNetGrad");
    if (!rc && hipMemcpy(grad_out, g_net->grad_, (size_t)g_net->nparam * 4,
                         hipMemcpyDeviceToHost) != hipSuccess)
        rc = -1;
    if (df) (void)hipFree(df);
    if (dt) (void)hipFree(dt);
    if (dm) (void)hipFree(dm);
    return rc;
}

API int seg4dNetForward(int I, int J, int K, int L, int P, const uint16_t* feat_bf16,
                        float* out) {
    if (!need_net("seg4dNetForward")) return -1;
    const size_t pts = (size_t)I * J * K * L;
    void* df = nullptr;
    float* dout = nullptr;
    int rc = 0;
    if (hipMalloc(&df, pts * 16) != hipSuccess || hipMalloc((void**)&dout, pts * 4) != hipSuccess ||
        hipMemcpy(df, feat_bf16, pts * 16, hipMemcpyHostToDevice) != hipSuccess) {
        s4d::set_error("seg4dNetForward: device buffers");
        rc = -1;
    }
    if (!rc && g_net->forward_batch(I, J, K, L, P, df, dout)) rc = net_fail("seg4dNetForward");
    if (!rc && hipMemcpy(out, dout, pts * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
    if (df) (void)hipFree(df);
    if (dout) (void)hipFree(dout);
    return rc;
}

API int seg4dNetPreview(float* c, double* q) {
    if (!need_net("seg4dNetPreview") || !need_engine("seg4dNetPreview")) return -1;
    if (!c || !q) {
        s4d::set_error("seg4dNetPreview: null output");
        return -1;
    }
    g_eng->net_ = g_net;
    if (g_eng->learned_preview(c, q)) {
        fail("seg4dNetPreview");
        return -1;
    }
    return 0;
}

API double seg4dNetPredictMs(void) { return g_net ? g_net->predict_ms : 0.0; }
API long seg4dNetPredicts(void) { return g_net ? g_net->predicts : 0; }
API void seg4dNetScales(float* s_d, float* s_t) {
    if (s_d) *s_d = g_net ? g_net->s_d : 0.f;
    if (s_t) *s_t = g_net ? g_net->s_t : 0.f;
}

API int seg4dSetSweepVariant(int v) {
    if (!g_eng) return -1;
    const int old = g_eng->sweep_variant_;
    g_eng->sweep_variant_ = (v >= 0 && v <= 63 && (v & 24) != 24) ? v : 0;
    return old;
}

API int seg4dSetSweepChunk(int lsteps) {
    if (!g_eng) return -1;
    const int old = g_eng->sweep_lc_;
    g_eng->sweep_lc_ = lsteps > 0 ? lsteps : 0;
    return old;
}

API int seg4dSetSweepBands(int nband) {
    if (!g_eng) return -1;
    const int old = g_eng->sweep_nband_;
    g_eng->sweep_nband_ = (nband > 0 && nband <= 64) ? nband : 0;
    return old;
}

API int seg4dSetPipelineDepth(int depth) {
    if (!g_eng) return -1;
    const int old = g_eng->pipe_depth_;
    g_eng->pipe_depth_ = (depth > 0 && depth <= 16) ? depth : 0;
    return old;
}

API int seg4dSetFusedSweeps(int mode) {
    if (!g_eng) return -2;
    const int old = g_eng->fused_;
    g_eng->fused_ = mode > 0 ? 1 : 0;
    return old;
}

API int seg4dLastSolveFused(void) { return g_eng && g_eng->fused_last_ ? 1 : 0; }

API int seg4dSetSweepTile(int runs, int planes) {
    if (!g_eng) return -1;
    if (runs < 1 || planes < 1) return -1;
    g_eng->sweep_tb_ = runs;
    g_eng->sweep_tk_ = planes;
    return 0;
}

API int seg4dL2normNoAlloc(const double* a, const double* b, long n, double hh, double* result,
                           void* stream) {
    // no engine, host array or device allocation is involved
    return s4d::l2norm_noalloc(a, b, n, hh, result, static_cast<hipStream_t>(stream));
}

API int seg4dInjectStepFailure(int where) {
    if (!g_eng) return -1;
    g_eng->inject_ = (where == 1 || where == 2) ? where : 0;
    return 0;
}

API bool seg4dSynchronize(void) {
    if (!g_eng) return true;
    return g_eng->sync() == 0;
}

API int seg4dDeviceInfo(char* buf, int len) {
    hipDeviceProp_t prop;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return -1;
    std::snprintf(buf, (size_t)len, "%s|%d CUs|%s|%.1f GB", prop.name, prop.multiProcessorCount,
                  prop.gcnArchName, prop.totalGlobalMem / 1e9);
    return 0;
}

// ---- single-volume launch-order autotuner ----------------------------------

API int seg4dAutotune(int reps, int* runs, int* planes, float* best_ms) {
    if (!need_engine("seg4dAutotune")) return -1;
    float ms = 0.f;
    if (g_eng->autotune(reps > 0 ? reps : 3, &ms)) {
        fail("seg4dAutotune");
        return -1;
    }
    const bool tiled = g_eng->tiled_order();
    if (runs) *runs = tiled ? g_eng->sweep_tb_ : 0;
    if (planes) *planes = tiled ? g_eng->sweep_tk_ : 0;
    if (best_ms) *best_ms = ms;
    return 0;
}

// ---- batched form -------------------------------------------------------------

static bool need_batch(const char* fn) {
    if (g_batch && g_batch->size() > 0) return true;
    s4d::set_error(std::string(fn) + ": no batch (seg4dBatchCreate first)");
    return false;
}

static int batch_fail(const char* fn) {
    std::string msg = fn;
    if (g_batch && !g_batch->err.empty()) msg += " (" + g_batch->err + ")";
    s4d::set_error(msg);
    std::fprintf(stderr, "seg4d: %s\n", msg.c_str());
    return -1;
}

API int seg4dBatchCreate(int nvol, const int* dims) {
    delete g_batch;
    g_batch = new s4d::Batch();
    if (g_batch->init(nvol, dims)) {
        const int rc = batch_fail("seg4dBatchCreate");
        delete g_batch;
        g_batch = nullptr;
        return rc;
    }
    return 0;
}

API void seg4dBatchDestroy(void) {
    delete g_batch;
    g_batch = nullptr;
}

API int seg4dBatchSetVolume(int v, const double* img, const double* u0, const double* mask) {
    if (!need_batch("seg4dBatchSetVolume")) return -1;
    s4d::EdgeParams ep;
    ep.sigma = sigma;
    ep.h = h;
    ep.omega = omega_sigma;
    ep.gsTol = gsTol;
    ep.rel2 = g_rel * g_rel;
    ep.cap = maxNoGSIteration_sigma;
    ep.K = K;
    ep.theta = theta_img;
    if (g_batch->set_volume(v, img, u0, mask, ep)) return batch_fail("seg4dBatchSetVolume");
    return 0;
}

API int seg4dBatchStep(void) {
    if (!need_batch("seg4dBatchStep")) return -1;
    if (g_batch->step(step_params())) return batch_fail("seg4dBatchStep");
    return 0;
}

API int seg4dBatchGetU(int v, double* out) {
    if (!need_batch("seg4dBatchGetU")) return -1;
    if (g_batch->download_u(v, out)) return batch_fail("seg4dBatchGetU");
    return 0;
}

API int seg4dBatchStats(int v, int* sweeps, int* rounds, double* R, double* tol) {
    if (!need_batch("seg4dBatchStats") || v < 0 || v >= g_batch->size()) return -1;
    const s4d::SolveStats& st = g_batch->stats[v];
    if (sweeps) *sweeps = st.z;
    if (rounds) *rounds = st.rounds;
    if (R) *R = st.R;
    if (tol) *tol = st.tol;
    return 0;
}

API double seg4dBatchLastStepMs(void) { return g_batch ? g_batch->last_step_ms : 0.0; }

API int seg4dBatchConfigure(int precision, int warm, int graph, int depth) {
    if (!need_batch("seg4dBatchConfigure")) return -1;
    if (precision < 0 || precision > 1 || !(warm == 0 || warm == 1 || warm == 3 || warm == 4) ||
        graph < 0 || graph > 2 || depth < 1 || depth > 16) {
        s4d::set_error("seg4dBatchConfigure: precision 0/1, warm 0/1/3/4, graph 0/1/2, depth 1..16");
        return -1;
    }
    g_batch->precision = precision;
    g_batch->warm = warm;
    g_batch->graph = graph;
    g_batch->depth = depth;
    return 0;
}

API int seg4dBatchGraphInfo(int* iters, int* rounds, long* steps, long* redone,
                            double* host_ms) {
    if (!need_batch("seg4dBatchGraphInfo")) return -1;
    if (host_ms) *host_ms = g_batch->graph_host_ms;
    if (iters) *iters = g_batch->graph_iters;
    if (rounds) *rounds = g_batch->graph_rounds;
    if (steps) *steps = g_batch->graph_steps;
    if (redone) *redone = g_batch->graph_redone;
    return 0;
}

API int seg4dBatchSetGraphPlan(int iters, int rounds) {
    if (!need_batch("seg4dBatchSetGraphPlan")) return -1;
    if (iters < 0 || rounds < 0 || (iters > 0) != (rounds > 0)) {
        s4d::set_error("seg4dBatchSetGraphPlan: iters and rounds both >= 1, or both 0");
        return -1;
    }
    g_batch->graph_iters = iters;
    g_batch->graph_rounds = rounds;
    return 0;
}

API int seg4dBatchSetStreams(int nstreams) {
    if (!need_batch("seg4dBatchSetStreams")) return -1;
    if (nstreams < 0 || nstreams > 64) {
        s4d::set_error("seg4dBatchSetStreams: 0 (one launch per pass) or 1..64 groups");
        return -1;
    }
    g_batch->streams = nstreams;
    return 0;
}

API int seg4dBatchAutotune(int reps, float* best_ms) {
    if (!need_batch("seg4dBatchAutotune")) return -1;
    std::vector<float> ms;
    if (g_batch->autotune(reps > 0 ? reps : 3, &ms)) return batch_fail("seg4dBatchAutotune");
    if (best_ms)
        for (int v = 0; v < g_batch->size(); v++) best_ms[v] = ms[v];
    return 0;
}

API int seg4dBatchTile(int v, int* runs, int* planes) {
    if (!need_batch("seg4dBatchTile") || v < 0 || v >= g_batch->size()) return -1;
    s4d::Engine& E = g_batch->vol(v);
    const bool tiled = E.tiled_order();
    if (runs) *runs = tiled ? E.sweep_tb_ : 0;
    if (planes) *planes = tiled ? E.sweep_tk_ : 0;
    return 0;
}
