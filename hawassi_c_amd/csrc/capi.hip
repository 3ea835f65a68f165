// capi.hip -- the extern "C" entry points declared in include/hawassi_hip.h.
#include <hip/hip_runtime.h>
#include <string>
#include <vector>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "solver2d.h"

namespace hw {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

std::vector<Entry>& registry() {
    static std::vector<Entry> r;
    return r;
}

// the full-grid path (full2d.hip)
Solver2D* make_solver_full(int precision);
bool full_supported(int nx, int ny, int precision);
int fft2_full_forward(int nx, int ny, int prec, const double* phys, double* re, double* im);
int fft2_full_inverse(int nx, int ny, int prec, const double* re, const double* im, double* phys, double* nyq_im);

static const Entry* find_entry(int prec, int nx, int ny) {
    for (const Entry& e : registry())
        if (e.prec == prec && e.nx == nx && e.ny == ny) return &e;
    return nullptr;
}

static int check_device() {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n < 1) {
        set_error("no HIP device available (%s)", e == hipSuccess ? "count 0" : hipGetErrorString(e));
        return HW_ERR_NOGPU;
    }
    return HW_OK;
}

// ---- scratch of the zero-allocation entry point (hw_fft2_forward_noalloc) --
// One array per GPU, held until the process ends.  Host code never uses its
// name as an address: each device's copy lives at its own address, which
// hipGetSymbolAddress gives once per device (na_scratch_ptr), and that
// pointer is what the kernels receive.
static __device__ __align__(256) unsigned char scratch[HW_NOALLOC_SCRATCH_BYTES];

static std::mutex g_na_mutex;
static void* g_na_ptr[64] = {};

static int na_scratch_ptr(unsigned char** out) {
    int dev = 0;
    HW_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) { set_error("device ordinal %d out of range", dev); return HW_ERR_ARG; }
    std::lock_guard<std::mutex> lk(g_na_mutex);
    if (!g_na_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
            return HW_ERR_HIP;
        }
        g_na_ptr[dev] = p;
    }
    *out = static_cast<unsigned char*>(g_na_ptr[dev]);
    return HW_OK;
}

}  // namespace hw

using namespace hw;

extern "C" {

const char* hw_last_error(void) { return g_err; }
const char* hw_version(void) { return "hawassi_hip 0.1 (gfx950)"; }

int hw_supported(int nx, int ny, int precision) { return find_entry(precision, nx, ny) ? 1 : 0; }

int hw_supported_full(int nx, int ny, int precision) {
    return (precision == HW_PREC_FP64 || precision == HW_PREC_FP32) && full_supported(nx, ny, precision) ? 1 : 0;
}

int hw_ctx_create_2d(const hw_config_2d* cfg, void** ctx) {
    if (!cfg || !ctx) { set_error("null argument"); return HW_ERR_ARG; }
    *ctx = nullptr;
    const Entry* e = find_entry(cfg->precision, cfg->nx, cfg->ny);
    // the fused path has no kernels for this grid, the caller asks for the
    // full grid, or the bottom is flat with walls: the full-grid path
    const bool full = cfg->full_grid || !e || (cfg->flat && cfg->wall);
    if (full && !hw_supported_full(cfg->nx, cfg->ny, cfg->precision)) {
        set_error("unsupported grid %dx%d precision %d (the full-grid path takes sides that factor into 2, 3 and 5, "
                  "at most 4096 in FP64 and 8192 in FP32)", cfg->nx, cfg->ny, cfg->precision);
        return HW_ERR_ARG;
    }
    int rc = check_device();
    if (rc) return rc;
    Solver2D* s = full ? make_solver_full(cfg->precision) : e->make();
    rc = s->init(cfg);
    if (rc) { delete s; return rc; }
    *ctx = s;
    return HW_OK;
}

int hw_ctx_create_2d_members(const hw_config_2d* cfg, int members, void** ctx) {
    if (!cfg || !ctx || members < 1) { set_error("bad argument"); return HW_ERR_ARG; }
    *ctx = nullptr;
    if (cfg->full_grid || (cfg->flat && cfg->wall)) {
        set_error("contexts with several members run on the fused path (full_grid = 0, no walls on a flat bottom)");
        return HW_ERR_ARG;
    }
    const Entry* e = find_entry(cfg->precision, cfg->nx, cfg->ny);
    if (!e) {
        set_error("unsupported grid %dx%d precision %d (no compiled kernels)", cfg->nx, cfg->ny, cfg->precision);
        return HW_ERR_ARG;
    }
    int rc = check_device();
    if (rc) return rc;
    Solver2D* s = e->make();
    s->nb = members;
    s->step_mode = HW_STEP_GRAPH;
    rc = s->init(cfg);
    if (rc) { delete s; return rc; }
    *ctx = s;
    return HW_OK;
}

void hw_ctx_destroy(void* ctx) { delete static_cast<Solver2D*>(ctx); }

// entry points that act on one state
static int one_member(const Solver2D* s, const char* fn) {
    if (s->nb == 1) return HW_OK;
    set_error("%s: the context holds %d members (hw_ctx_create_2d_members); use hw_set_state, hw_get_state and "
              "hw_integrate_members", fn, s->nb);
    return HW_ERR_ARG;
}

static int ensure_staging(Solver2D* s) {
    size_t bytes = sizeof(double) * 6 * (size_t)s->nx * s->ny;
    if (!s->d_full_a) HW_TRY(hipMalloc((void**)&s->d_full_a, bytes));
    if (!s->d_full_b) HW_TRY(hipMalloc((void**)&s->d_full_b, bytes));
    return HW_OK;
}

int hw_rhs(double t, const double* y, double* dydt, void* ctx) {
    Solver2D* s = static_cast<Solver2D*>(ctx);
    if (!s || !y || !dydt) { set_error("null argument"); return HW_ERR_ARG; }
    int rc = one_member(s, "hw_rhs");
    if (rc) return rc;
    rc = ensure_staging(s);
    if (rc) return rc;
    size_t bytes = sizeof(double) * 6 * (size_t)s->nx * s->ny;
    HW_TRY(hipMemcpyAsync(s->d_full_a, y, bytes, hipMemcpyHostToDevice, s->stream));
    rc = s->rhs_full_dev(t, s->d_full_a, s->d_full_b);
    if (rc) return rc;
    HW_TRY(hipMemcpyAsync(dydt, s->d_full_b, bytes, hipMemcpyDeviceToHost, s->stream));
    HW_TRY(hipStreamSynchronize(s->stream));
    return HW_OK;
}

int hw_rhs_device(void* ctx, double t, const double* d_y, double* d_dydt) {
    Solver2D* s = static_cast<Solver2D*>(ctx);
    if (!s || !d_y || !d_dydt) { set_error("null argument"); return HW_ERR_ARG; }
    int rc = one_member(s, "hw_rhs_device");
    if (rc) return rc;
    return s->rhs_full_dev(t, d_y, d_dydt);
}

int hw_set_state(void* ctx, const double* y) {
    Solver2D* s = static_cast<Solver2D*>(ctx);
    if (!s || !y) { set_error("null argument"); return HW_ERR_ARG; }
    int rc = ensure_staging(s);
    if (rc) return rc;
    size_t n6 = 6 * (size_t)s->nx * s->ny, bytes = sizeof(double) * n6;
    if (s->nb > 1) {
        // one full state after the other
        for (int z = 0; z < s->nb; ++z) {
            HW_TRY(hipMemcpyAsync(s->d_full_a, y + z * n6, bytes, hipMemcpyHostToDevice, s->stream));
            if ((rc = s->set_member_dev(s->d_full_a, z))) return rc;
        }
        return HW_OK;
    }
    HW_TRY(hipMemcpyAsync(s->d_full_a, y, bytes, hipMemcpyHostToDevice, s->stream));
    return s->set_state_dev(s->d_full_a);
}

int hw_get_state(void* ctx, double* y) {
    Solver2D* s = static_cast<Solver2D*>(ctx);
    if (!s || !y) { set_error("null argument"); return HW_ERR_ARG; }
    int rc = ensure_staging(s);
    if (rc) return rc;
    size_t n6 = 6 * (size_t)s->nx * s->ny, bytes = sizeof(double) * n6;
    if (s->nb > 1) {
        for (int z = 0; z < s->nb; ++z) {
            if ((rc = s->get_member_dev(s->d_full_a, z))) return rc;
            HW_TRY(hipMemcpyAsync(y + z * n6, s->d_full_a, bytes, hipMemcpyDeviceToHost, s->stream));
            HW_TRY(hipStreamSynchronize(s->stream));
        }
        return HW_OK;
    }
    rc = s->get_state_dev(s->d_full_a);
    if (rc) return rc;
    HW_TRY(hipMemcpyAsync(y, s->d_full_a, bytes, hipMemcpyDeviceToHost, s->stream));
    HW_TRY(hipStreamSynchronize(s->stream));
    return HW_OK;
}

int hw_evolve_apply(void* ctx, double* t, double t1, double* h) {
    Solver2D* s = static_cast<Solver2D*>(ctx);
    if (!s || !t || !h) { set_error("null argument"); return HW_ERR_ARG; }
    int rc = one_member(s, "hw_evolve_apply");
    if (rc) return rc;
    return s->evolve_apply(t, t1, h);
}

int hw_set_step_mode(void* ctx, int mode) {
    Solver2D* s = static_cast<Solver2D*>(ctx);
    if (!s || (mode != HW_STEP_HOST && mode != HW_STEP_GRAPH)) { set_error("bad argument"); return HW_ERR_ARG; }
    if (s->nb > 1 && mode != HW_STEP_GRAPH) {
        set_error("a context with several members steps on the device only");
        return HW_ERR_ARG;
    }
    s->step_mode = mode;
    return HW_OK;
}

// t and h hold one value per member of every context, context after context
static int graph_batch(Solver2D* const* ss, int n, double* t, double t1, double* h) {
    std::vector<char> act((size_t)n, 0);
    std::vector<size_t> off((size_t)n + 1, 0);
    int rc;
    for (int i = 0; i < n; ++i) off[i + 1] = off[i] + (size_t)ss[i]->nb;
    for (int i = 0; i < n; ++i) {
        bool any = false;
        for (size_t j = off[i]; j < off[i + 1]; ++j) {
            const double dt = t1 - t[j];
            if ((dt < 0.0 && h[j] > 0.0) || (dt > 0.0 && h[j] < 0.0)) {
                set_error("context %d: step direction must match interval direction", i);
                return HW_ERR_ARG;
            }
            any = any || t[j] < t1;
        }
        if (!any) continue;
        if ((rc = ss[i]->graph_begin(t + off[i], t1, h + off[i]))) return rc;
        act[i] = 1;
    }
    for (;;) {
        bool any = false;
        // every unfinished context's next batch of attempts, each on its own stream ...
        for (int i = 0; i < n; ++i)
            if (act[i]) {
                any = true;
                if ((rc = ss[i]->graph_launch())) return rc;
            }
        if (!any) break;
        // ... then each one's progress
        for (int i = 0; i < n; ++i) {
            if (!act[i]) continue;
            bool done = false;
            if ((rc = ss[i]->graph_poll(&done))) return rc;
            if (done) {
                act[i] = 0;
                if ((rc = ss[i]->graph_finish(t + off[i], h + off[i]))) {
                    const std::string m = hw_last_error();
                    set_error("context %d: %s", i, m.c_str());
                    return rc;
                }
            }
        }
    }
    return HW_OK;
}

int hw_integrate_members(void* ctx, double* t, double t1, double* h) {
    Solver2D* s = static_cast<Solver2D*>(ctx);
    if (!s || !t || !h) { set_error("null argument"); return HW_ERR_ARG; }
    if (s->step_mode != HW_STEP_GRAPH) { set_error("hw_integrate_members needs HW_STEP_GRAPH"); return HW_ERR_ARG; }
    return graph_batch(&s, 1, t, t1, h);
}

int hw_integrate(void* ctx, double* t, double t1, double* h) {
    Solver2D* s = static_cast<Solver2D*>(ctx);
    if (!s || !t || !h) { set_error("null argument"); return HW_ERR_ARG; }
    int rc = one_member(s, "hw_integrate");
    if (rc) return rc;
    if (s->step_mode == HW_STEP_GRAPH) return graph_batch(&s, 1, t, t1, h);
    while (*t < t1) {
        int rc = s->evolve_apply(t, t1, h);
        if (rc) return rc;
    }
    return HW_OK;
}

int hw_integrate_batch(void* const* ctxs, int n, double* t, double t1, double* h) {
    if (n < 0 || (n > 0 && (!ctxs || !t || !h))) { set_error("null argument"); return HW_ERR_ARG; }
    {
        bool all_graph = n > 0, multi = false;
        for (int i = 0; i < n; ++i) {
            if (!ctxs[i]) { set_error("null context %d", i); return HW_ERR_ARG; }
            all_graph = all_graph && static_cast<Solver2D*>(ctxs[i])->step_mode == HW_STEP_GRAPH;
            multi = multi || static_cast<Solver2D*>(ctxs[i])->nb > 1;
        }
        // contexts holding several members take part only when every context
        // steps on the device; t and h then hold one value per member
        if (multi && !all_graph) {
            for (int i = 0; i < n; ++i) {
                int rc = one_member(static_cast<Solver2D*>(ctxs[i]), "hw_integrate_batch");
                if (rc) return rc;
            }
        }
        if (all_graph) {
            std::vector<Solver2D*> ss((size_t)n);
            for (int i = 0; i < n; ++i) ss[i] = static_cast<Solver2D*>(ctxs[i]);
            return graph_batch(ss.data(), n, t, t1, h);
        }
    }
    struct Lane { double t0, h0; bool fin, need_k0, active; };
    std::vector<Lane> ln((size_t)n);
    for (int i = 0; i < n; ++i) {
        if (!ctxs[i]) { set_error("null context %d", i); return HW_ERR_ARG; }
        const double dt = t1 - t[i];
        if ((dt < 0.0 && h[i] > 0.0) || (dt > 0.0 && h[i] < 0.0)) {
            set_error("context %d: step direction must match interval direction", i);
            return HW_ERR_ARG;
        }
        ln[i] = Lane{t[i], h[i], false, true, t[i] < t1};
    }
    for (;;) {
        bool any = false;
        // enqueue one attempt per unfinished context (each on its own stream) ...
        for (int i = 0; i < n; ++i) {
            Lane& L = ln[i];
            if (!L.active) continue;
            any = true;
            const double dt = t1 - L.t0;
            L.fin = false;
            if ((dt >= 0.0 && L.h0 > dt) || (dt < 0.0 && L.h0 < dt)) { L.h0 = dt; L.fin = true; }
            int rc = static_cast<Solver2D*>(ctxs[i])->attempt_issue(L.t0, L.h0, L.need_k0);
            if (rc) return rc;
            L.need_k0 = false;
        }
        if (!any) return HW_OK;
        // ... then apply each one's step control (gsl_odeiv2_evolve_apply + control_y)
        for (int i = 0; i < n; ++i) {
            Lane& L = ln[i];
            if (!L.active) continue;
            Solver2D* s = static_cast<Solver2D*>(ctxs[i]);
            double rm;
            int rc = s->attempt_result(&rm);
            if (rc) return rc;
            const double tn = L.fin ? t1 : L.t0 + L.h0;
            const double h_old = L.h0;
            double hn = L.h0;
            if (hadjust_rule(rm, &hn) < 0) {
                if (std::fabs(hn) < std::fabs(h_old) && tn + hn != tn) {
                    s->st.steps_rejected++;
                    L.h0 = hn;
                    continue;
                }
                t[i] = L.t0;
                h[i] = hn;
                set_error("context %d: step size underflow at t=%g", i, L.t0);
                return HW_ERR_STEP;
            }
            if ((rc = s->accept_step(tn))) return rc;
            t[i] = L.t0 = tn;
            h[i] = L.h0 = hn;
            L.need_k0 = true;
            L.active = tn < t1;
        }
    }
}

int hw_get_fields(void* ctx, double* eta, double* u, double* v) {
    Solver2D* s = static_cast<Solver2D*>(ctx);
    if (!s) { set_error("null argument"); return HW_ERR_ARG; }
    size_t n = (size_t)s->nx * s->ny;
    int rc = one_member(s, "hw_get_fields");
    if (rc) return rc;
    rc = ensure_staging(s);
    if (rc) return rc;
    double* d[3] = {eta ? s->d_full_a : nullptr, u ? s->d_full_a + n : nullptr, v ? s->d_full_a + 2 * n : nullptr};
    rc = s->fields_dev(d[0], d[1], d[2]);
    if (rc) return rc;
    double* h[3] = {eta, u, v};
    for (int f = 0; f < 3; ++f)
        if (h[f]) HW_TRY(hipMemcpy(h[f], d[f], sizeof(double) * n, hipMemcpyDeviceToHost));
    return HW_OK;
}

int hw_set_tprev(void* ctx, double tprev) {
    Solver2D* s = static_cast<Solver2D*>(ctx);
    if (!s) { set_error("null argument"); return HW_ERR_ARG; }
    s->tprev = tprev;
    return HW_OK;
}

int hw_get_breaking(void* ctx, double* B, int* flagbr) {
    Solver2D* s = static_cast<Solver2D*>(ctx);
    if (!s) { set_error("null argument"); return HW_ERR_ARG; }
    return s->breaking_state(B, flagbr);
}

int hw_get_stats(void* ctx, hw_stats* st) {
    Solver2D* s = static_cast<Solver2D*>(ctx);
    if (!s || !st) { set_error("null argument"); return HW_ERR_ARG; }
    *st = s->st;
    return HW_OK;
}

int hw_reset_stats(void* ctx) {
    Solver2D* s = static_cast<Solver2D*>(ctx);
    if (!s) { set_error("null argument"); return HW_ERR_ARG; }
    s->st = hw_stats{};
    return HW_OK;
}

int hw_sync(void* ctx) {
    Solver2D* s = static_cast<Solver2D*>(ctx);
    if (!s) { set_error("null argument"); return HW_ERR_ARG; }
    HW_TRY(hipStreamSynchronize(s->stream));
    return HW_OK;
}

int hw_fft2_forward(int nx, int ny, int precision, const double* phys, double* re, double* im) {
    const Entry* e = find_entry(precision, nx, ny);
    if (!e && !hw_supported_full(nx, ny, precision)) {
        set_error("unsupported grid %dx%d precision %d", nx, ny, precision);
        return HW_ERR_ARG;
    }
    if (!phys || !re || !im) { set_error("null argument"); return HW_ERR_ARG; }
    int rc = check_device();
    if (rc) return rc;
    return e ? e->fwd(phys, re, im) : fft2_full_forward(nx, ny, precision, phys, re, im);
}

int hw_fft2_inverse(int nx, int ny, int precision, const double* re, const double* im, double* phys, double* nyq_im) {
    const Entry* e = find_entry(precision, nx, ny);
    if (!e && !hw_supported_full(nx, ny, precision)) {
        set_error("unsupported grid %dx%d precision %d", nx, ny, precision);
        return HW_ERR_ARG;
    }
    if (!phys || !re || !im) { set_error("null argument"); return HW_ERR_ARG; }
    int rc = check_device();
    if (rc) return rc;
    return e ? e->inv(re, im, phys, nyq_im) : fft2_full_inverse(nx, ny, precision, re, im, phys, nyq_im);
}

int hw_fft2_forward_noalloc(int nx, int ny, int precision, const double* d_phys, double* d_re, double* d_im,
                            void* stream) {
    const Entry* e = find_entry(precision, nx, ny);
    if (!e) { set_error("unsupported grid %dx%d precision %d", nx, ny, precision); return HW_ERR_ARG; }
    if (!d_phys || !d_re || !d_im) { set_error("null argument"); return HW_ERR_ARG; }
    unsigned char* ws = nullptr;
    int rc = na_scratch_ptr(&ws);
    if (rc) return rc;
    return e->fwd_na(d_phys, d_re, d_im, static_cast<hipStream_t>(stream), ws, (size_t)HW_NOALLOC_SCRATCH_BYTES);
}

}  // extern "C"
