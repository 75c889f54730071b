// Error reporting, timers and small host utilities.
//
// The last-error record is thread-local and keeps module / file / line /
// reason of the most recent failure, with the reference's accessors
// (src/helper_funcs/error.c:27-118).  Unlike the reference it is recorded in
// every build type, and messages go to stderr only when HIPARTI_VERBOSE is
// set, so library users see errors through return codes.
#include "internal.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>

namespace {

struct LastError {
    const char *module = nullptr;
    int errcode = 0;
    const char *file = nullptr;
    unsigned line = 0;
    std::string reason;
    bool has_reason = false;
};

thread_local LastError g_err;

}  // namespace

namespace pti {

bool verbose() {
    static int v = -1;
    if (v < 0) {
        const char *s = std::getenv("HIPARTI_VERBOSE");
        v = (s && s[0] && s[0] != '0') ? 1 : 0;
    }
    return v == 1;
}

int default_threads() {
    const char *s = std::getenv("OMP_NUM_THREADS");
    if (s) {
        int n = std::atoi(s);
        if (n > 0) return n;
    }
    unsigned hc = std::thread::hardware_concurrency();
    return hc ? (int)hc : 1;
}

void parallel_for(uint64_t n, int nthreads,
                  const std::function<void(uint64_t, uint64_t, int)> &body) {
    if (nthreads <= 0) nthreads = default_threads();
    if ((uint64_t)nthreads > n) nthreads = (int)std::max<uint64_t>(n, 1);
    if (nthreads == 1) {
        body(0, n, 0);
        return;
    }
    std::vector<std::thread> ts;
    ts.reserve(nthreads);
    for (int t = 0; t < nthreads; ++t) {
        uint64_t b = n * (uint64_t)t / (uint64_t)nthreads;
        uint64_t e = n * (uint64_t)(t + 1) / (uint64_t)nthreads;
        ts.emplace_back([&body, b, e, t] { body(b, e, t); });
    }
    for (auto &th : ts) th.join();
}

}  // namespace pti

extern "C" {

void pti_ComplainError(const char *module, int errcode, const char *file, unsigned line,
                       const char *reason) {
    g_err.module = module;
    g_err.errcode = errcode;
    g_err.file = file;
    g_err.line = line;
    g_err.has_reason = reason != nullptr;
    g_err.reason = reason ? reason : "";
    if (pti::verbose()) {
        std::fprintf(stderr, "[%s] error 0x%08x at %s:%u%s%s\n", module ? module : "?", errcode,
                     file ? file : "?", line, reason ? ", " : "", reason ? reason : "");
    }
}

void pti_Panic(const char *file, unsigned line, const char *expr) {
    std::fprintf(stderr, "%s:%u: assertion error: \"%s\"\n", file, line, expr);
    std::abort();
}

int ptiGetLastError(const char **module, const char **file, unsigned *line, const char **reason) {
    if (module) *module = g_err.module;
    if (file) *file = g_err.file;
    if (line) *line = g_err.line;
    if (reason) *reason = g_err.has_reason ? g_err.reason.c_str() : nullptr;
    return g_err.errcode;
}

void ptiClearLastError(void) {
    g_err = LastError();
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Host wall-clock timer.  A device-event timer is created when use_cuda != 0;
// that variant lives in the device sources (timer_dev.hip) and is reached
// through these hooks so host code does not depend on HIP headers.
// ---------------------------------------------------------------------------
struct ptiTagTimer {
    int use_dev;
    std::chrono::steady_clock::time_point t0, t1;
    void *dev;  // device-event state, owned by timer_dev.hip
};

namespace pti {
int dev_timer_create(void **state);
int dev_timer_start(void *state);
int dev_timer_stop(void *state);
double dev_timer_elapsed(void *state);
void dev_timer_destroy(void *state);
}  // namespace pti

extern "C" {

int ptiNewTimer(ptiTimer *timer, int use_cuda) {
    ptiTagTimer *t = new (std::nothrow) ptiTagTimer();
    PTI_CHECK_ALLOC(t, "Timer New");
    t->use_dev = use_cuda != 0;
    t->dev = nullptr;
    if (t->use_dev) {
        int rc = pti::dev_timer_create(&t->dev);
        if (rc) {
            delete t;
            PTI_FAIL(rc, "Timer New", "device timer");
        }
    }
    *timer = t;
    return 0;
}

int ptiStartTimer(ptiTimer timer) {
    if (timer->use_dev) return pti::dev_timer_start(timer->dev);
    timer->t0 = std::chrono::steady_clock::now();
    return 0;
}

int ptiStopTimer(ptiTimer timer) {
    if (timer->use_dev) return pti::dev_timer_stop(timer->dev);
    timer->t1 = std::chrono::steady_clock::now();
    return 0;
}

double ptiElapsedTime(const ptiTimer timer) {
    if (timer->use_dev) return pti::dev_timer_elapsed(timer->dev);
    return std::chrono::duration<double>(timer->t1 - timer->t0).count();
}

double ptiPrintElapsedTime(const ptiTimer timer, const char *name) {
    double t = ptiElapsedTime(timer);
    std::printf("[%s]: %.9lf s\n", name, t);
    return t;
}

double ptiPrintAverageElapsedTime(const ptiTimer timer, const int niters, const char *name) {
    double t = ptiElapsedTime(timer) / niters;
    std::printf("[%s]: %.9lf s\n", name, t);
    return t;
}

int ptiFreeTimer(ptiTimer timer) {
    if (!timer) return 0;
    if (timer->use_dev) pti::dev_timer_destroy(timer->dev);
    delete timer;
    return 0;
}

}  // extern "C"
