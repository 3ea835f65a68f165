// capi1d.hip -- the extern "C" 1-D entry points declared in
// include/hawassi_hip.h and the launch dispatch over size classes.
#include <hip/hip_runtime.h>

#include "solver1d.h"

namespace hw {

// defined in the generated units: per (precision, size class n) the RHS
// launch and its LDS attribute, per (precision, thread count) the control
#define HW_1D_DECL_N(T, P, N)                                                                             \
    hipError_t k1d_launch_##P##_##N(dim3, size_t, hipStream_t, const Batch1DArgs<T>&, int, int, int);     \
    hipError_t k1d_setlds_##P##_##N(size_t);
#define HW_1D_DECL_NT(T, P, NT) hipError_t k1d_control_##P##_##NT(dim3, hipStream_t, const Batch1DArgs<T>&, int);
#define HW_1D_DECL_ALL(T, P) HW_1D_DECL_N(T, P, 128) HW_1D_DECL_N(T, P, 256) HW_1D_DECL_N(T, P, 512) \
    HW_1D_DECL_N(T, P, 1024) HW_1D_DECL_N(T, P, 2000) HW_1D_DECL_N(T, P, 2048) HW_1D_DECL_N(T, P, 4096) \
    HW_1D_DECL_N(T, P, 8192) HW_1D_DECL_NT(T, P, 64) HW_1D_DECL_NT(T, P, 128) HW_1D_DECL_NT(T, P, 256) \
    HW_1D_DECL_NT(T, P, 512) HW_1D_DECL_NT(T, P, 1000) HW_1D_DECL_NT(T, P, 1024)
HW_1D_DECL_ALL(float, 1)
HW_1D_DECL_ALL(double, 0)
HW_1D_DECL_N(float, 1, 16384)

#define HW_1D_SWITCH_N(P, CALL, ...)                                        \
    switch (n) {                                                            \
        case 128: return CALL##_##P##_128(__VA_ARGS__);                     \
        case 256: return CALL##_##P##_256(__VA_ARGS__);                     \
        case 512: return CALL##_##P##_512(__VA_ARGS__);                     \
        case 1024: return CALL##_##P##_1024(__VA_ARGS__);                   \
        case 2000: return CALL##_##P##_2000(__VA_ARGS__);                   \
        case 2048: return CALL##_##P##_2048(__VA_ARGS__);                   \
        case 4096: return CALL##_##P##_4096(__VA_ARGS__);                   \
        case 8192: return CALL##_##P##_8192(__VA_ARGS__);                   \
        HW_1D_CASE_16384_##P(CALL, __VA_ARGS__)                             \
        default: return hipErrorInvalidValue;                               \
    }
#define HW_1D_CASE_16384_1(CALL, ...) case 16384: return CALL##_1_16384(__VA_ARGS__);
#define HW_1D_CASE_16384_0(CALL, ...)
#define HW_1D_SWITCH_NT(P, CALL, ...)                                       \
    switch (NT) {                                                           \
        case 64: return CALL##_##P##_64(__VA_ARGS__);                       \
        case 128: return CALL##_##P##_128(__VA_ARGS__);                     \
        case 256: return CALL##_##P##_256(__VA_ARGS__);                     \
        case 512: return CALL##_##P##_512(__VA_ARGS__);                     \
        case 1000: return CALL##_##P##_1000(__VA_ARGS__);                   \
        case 1024: return CALL##_##P##_1024(__VA_ARGS__);                   \
        default: return hipErrorInvalidValue;                               \
    }

template <> hipError_t K1D<float>::launch(int n, dim3 g, size_t lds, hipStream_t s, const Batch1DArgs<float>& a, int st,
                                          int list0, int mid) {
    HW_1D_SWITCH_N(1, k1d_launch, g, lds, s, a, st, list0, mid)
}
template <> hipError_t K1D<float>::control(int NT, dim3 g, hipStream_t s, const Batch1DArgs<float>& a, int list0) {
    HW_1D_SWITCH_NT(1, k1d_control, g, s, a, list0)
}
template <> hipError_t K1D<float>::set_lds(int n, size_t b) { HW_1D_SWITCH_N(1, k1d_setlds, b) }
template <> hipError_t K1D<double>::launch(int n, dim3 g, size_t lds, hipStream_t s, const Batch1DArgs<double>& a, int st,
                                           int list0, int mid) {
    HW_1D_SWITCH_N(0, k1d_launch, g, lds, s, a, st, list0, mid)
}
template <> hipError_t K1D<double>::control(int NT, dim3 g, hipStream_t s, const Batch1DArgs<double>& a, int list0) {
    HW_1D_SWITCH_NT(0, k1d_control, g, s, a, list0)
}
template <> hipError_t K1D<double>::set_lds(int n, size_t b) { HW_1D_SWITCH_N(0, k1d_setlds, b) }

}  // namespace hw

using namespace hw;

extern "C" {

int hw1d_create(const hw_config_1d* cfgs, int nreal, int precision, void** ctx) {
    if (!cfgs || !ctx || nreal < 1) { set_error("null argument or empty batch"); return HW_ERR_ARG; }
    *ctx = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev < 1) {
        set_error("no HIP device available (%s)", e == hipSuccess ? "count 0" : hipGetErrorString(e));
        return HW_ERR_NOGPU;
    }
    Solver1D* s = nullptr;
    if (precision == HW_PREC_FP32) s = new Solver1DT<float>();
    else if (precision == HW_PREC_FP64) s = new Solver1DT<double>();
    else { set_error("unknown precision %d", precision); return HW_ERR_ARG; }
    int rc = s->init(cfgs, nreal);
    if (rc) { delete s; return rc; }
    *ctx = s;
    return HW_OK;
}

void hw1d_destroy(void* ctx) { delete static_cast<Solver1D*>(ctx); }

int hw1d_rhs_at(void* ctx, int r, double t, const double* y, double* dydt) {
    Solver1D* s = static_cast<Solver1D*>(ctx);
    if (!s || !y || !dydt) { set_error("null argument"); return HW_ERR_ARG; }
    return s->rhs_at(r, t, y, dydt);
}

int hw1d_rhs(double t, const double* y, double* dydt, void* ctx) { return hw1d_rhs_at(ctx, 0, t, y, dydt); }

int hw1d_set_state(void* ctx, int r, const double* y) {
    Solver1D* s = static_cast<Solver1D*>(ctx);
    if (!s || !y) { set_error("null argument"); return HW_ERR_ARG; }
    return s->set_state(r, y);
}

int hw1d_get_state(void* ctx, int r, double* y) {
    Solver1D* s = static_cast<Solver1D*>(ctx);
    if (!s || !y) { set_error("null argument"); return HW_ERR_ARG; }
    return s->get_state(r, y);
}

int hw1d_set_time(void* ctx, int r, double t, double h) {
    Solver1D* s = static_cast<Solver1D*>(ctx);
    if (!s) { set_error("null argument"); return HW_ERR_ARG; }
    if (!(h > 0)) { set_error("step must be positive"); return HW_ERR_ARG; }
    return s->set_time(r, t, h);
}

int hw1d_get_time(void* ctx, int r, double* t, double* h) {
    Solver1D* s = static_cast<Solver1D*>(ctx);
    if (!s) { set_error("null argument"); return HW_ERR_ARG; }
    return s->get_time(r, t, h);
}

int hw1d_integrate(void* ctx, double t1) {
    Solver1D* s = static_cast<Solver1D*>(ctx);
    if (!s) { set_error("null argument"); return HW_ERR_ARG; }
    return s->integrate(t1, 0L);
}

int hw1d_set_round_limit(void* ctx, long max_rounds) {
    Solver1D* s = static_cast<Solver1D*>(ctx);
    if (!s || max_rounds < 0) { set_error("null context or negative limit"); return HW_ERR_ARG; }
    s->round_limit = max_rounds;
    return HW_OK;
}

int hw1d_get_breaking(void* ctx, int r, double* B, int* flag) {
    Solver1D* s = static_cast<Solver1D*>(ctx);
    if (!s) { set_error("null argument"); return HW_ERR_ARG; }
    return s->breaking_state(r, B, flag);
}

int hw1d_get_stats(void* ctx, int r, hw_stats* st) {
    Solver1D* s = static_cast<Solver1D*>(ctx);
    if (!s || !st) { set_error("null argument"); return HW_ERR_ARG; }
    if (r < -1 || r >= s->nreal) { set_error("realization %d out of range", r); return HW_ERR_ARG; }
    return s->stats(r, st);
}

}  // extern "C"
