// full2d.hip -- instantiations of the full-grid 2-D path (solver2dg.h) in
// FP64 and FP32; the C ABI (capi.hip) creates them through make_solver_full.
#include "solver2dg.h"

namespace hw {

Solver2D* make_solver_full(int precision) {
    if (precision == HW_PREC_FP64) return new Solver2DG<double>();
    if (precision == HW_PREC_FP32) return new Solver2DG<float>();
    return nullptr;
}

// whether the full-grid path can take an nx x ny grid in this precision
bool full_supported(int nx, int ny, int precision) {
    PlanG a, b;
    if (!plan_g(nx, a) || !plan_g(ny, b) || nx < 2 || ny < 2) return false;
    const size_t es = precision == HW_PREC_FP64 ? 16 : 8;
    return 2 * (size_t)nx * es <= Fft2G<double2>::LDS_MAX && 2 * (size_t)ny * es <= Fft2G<double2>::LDS_MAX;
}

// stand-alone transforms of the full-grid path (hw_fft2_forward /
// hw_fft2_inverse on grids without fused kernels): forward of a real field
// into the full spectrum; inverse of a full spectrum after the Hermitian
// projection, real part / (nx ny), and the imaginary part at x = 0 (the
// (-1)^x s(y) term of the projected inverse, s(y) at x = 0)
template <typename T>
static int fft2_full_t(int nx, int ny, int dir, const double* a_re, const double* a_im, double* o_re, double* o_im) {
    using C = typename ctype<T>::type;
    const size_t N = (size_t)nx * ny;
    Fft2G<C> f;
    if (const char* m = f.init(nx, ny, 0)) { set_error("%s", m); return HW_ERR_ARG; }
    std::vector<C> h(N);
    for (size_t q = 0; q < N; ++q) { h[q].x = (T)a_re[q]; h[q].y = a_im ? (T)a_im[q] : (T)0; }
    DevBuf<C> d;
    int rc = d.upload(h, 0);
    if (rc) return rc;
    if (dir > 0) k_proj_full<C><<<(unsigned)((N + 255) / 256), 256>>>(d.p, nx, ny);
    f.run(d.p, dir, 0);
    HW_TRY(hipGetLastError());
    HW_TRY(hipMemcpy(h.data(), d.p, sizeof(C) * N, hipMemcpyDeviceToHost));
    for (size_t q = 0; q < N; ++q) {
        o_re[q] = h[q].x;
        if (o_im) o_im[q] = h[q].y;
    }
    return HW_OK;
}

int fft2_full_forward(int nx, int ny, int prec, const double* phys, double* re, double* im) {
    return prec == HW_PREC_FP64 ? fft2_full_t<double>(nx, ny, -1, phys, nullptr, re, im)
                                : fft2_full_t<float>(nx, ny, -1, phys, nullptr, re, im);
}

int fft2_full_inverse(int nx, int ny, int prec, const double* re, const double* im, double* phys, double* nyq_im) {
    const size_t N = (size_t)nx * ny;
    std::vector<double> a(N), b(N);
    int rc = prec == HW_PREC_FP64 ? fft2_full_t<double>(nx, ny, 1, re, im, a.data(), b.data())
                                  : fft2_full_t<float>(nx, ny, 1, re, im, a.data(), b.data());
    if (rc) return rc;
    for (size_t q = 0; q < N; ++q) phys[q] = a[q] / (double)N;
    if (nyq_im)
        for (int i = 0; i < ny; ++i) nyq_im[i] = b[(size_t)i * nx] / (double)N;
    return HW_OK;
}

}  // namespace hw
