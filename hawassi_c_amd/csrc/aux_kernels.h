// aux_kernels.h -- RKF45 stage/error kernels, boundary layout conversion,
// influx-signal evaluation and stand-alone 2-D transforms.
#pragma once
#include "rhs2d_kernels.h"

namespace hw {

// ------------------------------------------------------------- RKF45 ------
// Fehlberg 4(5) coefficients as used by GSL's rkf45 stepper
// (propagates the 5th-order solution, error = 5th - 4th order).
struct StageCoef {
    const void* k[5];
    double b[5];
    int nk;
};

// The stage, final and norm arithmetic below is compiled without
// floating-point contraction: every product and sum is rounded on its own,
// as in the FP64 oracle, so the host-stepped kernels (k_rk_stage,
// k_rk_final) and the graph-stepped ones (k_rk_stage_g, k_rk_final_g) give
// the same bits whatever fused multiply-adds the compiler would pick in each.
// The error estimate is a small difference of the stage derivatives; left
// to the compiler, the two kernel shapes contracted differently and their
// step sizes parted by ~3e-11 per step.

// out = y + h * sum_j b_j k_j, converted to the RHS precision
template <typename T>
__device__ __forceinline__ void rk_stage_body(const double2* __restrict__ y, const StageCoef& sc, double h,
                                              typename ctype<T>::type* __restrict__ out, size_t n) {
#pragma clang fp contract(off)
    using C = typename ctype<T>::type;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        double2 yv = y[i];
        double ax = 0.0, ay = 0.0;
        for (int j = 0; j < sc.nk; ++j) {
            C kv = reinterpret_cast<const C*>(sc.k[j])[i];
            ax += sc.b[j] * (double)kv.x;
            ay += sc.b[j] * (double)kv.y;
        }
        double rx = yv.x + h * ax, ry = yv.y + h * ay;
        out[i] = cmk<C>((T)rx, (T)ry);
    }
}

template <typename T>
__global__ void k_rk_stage(const double2* __restrict__ y, StageCoef sc, double h, typename ctype<T>::type* __restrict__ out, size_t n) {
    rk_stage_body<T>(y, sc, h, out, n);
}

// control_y hadjust (GSL standard control with a_y = 1, a_dydt = 0, order 5):
// returns -1 decrease, 0 keep, +1 increase.  The host step loop and the
// device step control (k_step_end) share it.
__host__ __device__ inline int hadjust_rule(double rmax, double* h) {
    const double S = 0.9;
    const int ord = 5;
    double h_old = *h;
    if (rmax > 1.1) {
        double r = S / pow(rmax, 1.0 / ord);
        if (r < 0.2) r = 0.2;
        *h = r * h_old;
        return -1;
    } else if (rmax < 0.5) {
        double r = S / pow(rmax, 1.0 / (ord + 1.0));
        if (r > 5.0) r = 5.0;
        if (r < 1.0) r = 1.0;
        *h = r * h_old;
        return 1;
    }
    return 0;
}

// ------------------------------------------------- device step control ----
// gsl_odeiv2_evolve_apply's state kept in device memory, so that one
// attempted step -- K0, the five stages, the error norm and control_y -- is a
// fixed sequence of launches that a HIP graph replays without the host:
// k_step_begin clips the step to t1, the stage and final kernels take h and
// the buffer holding y from here, k_step_end accepts (swapping the buffers)
// or rejects.  Once t reaches t1 (or the step underflows) `done` is set and
// every kernel of a further replay returns at once.
struct DevStep {
    double t, h, t1;            // evolve_apply's t and h, the driver's target
    double t0, h0;              // this attempt's start and step
    double eps_abs, eps_rel;
    unsigned long long rmax_bits;
    int cur;                    // buffer holding y: 0 or 1
    int done;                   // t reached t1, or the step underflowed
    int err;                    // step size underflow at t0
    int final_step;
    long long accepted, rejected, attempts;
};

static __global__ void k_step_begin(DevStep* sv, int nb) {
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= nb) return;
    DevStep* s = sv + z;
    if (s->done) return;
    const double dt = s->t1 - s->t;
    double h0 = s->h;
    int fin = 0;
    if ((dt >= 0.0 && h0 > dt) || (dt < 0.0 && h0 < dt)) { h0 = dt; fin = 1; }
    s->t0 = s->t;
    s->h0 = h0;
    s->final_step = fin;
    s->rmax_bits = 0ull;
}

// members are consecutive ranges of `per` elements, each with its own h and
// state buffer; a finished member's range is left alone
template <typename T>
__global__ void k_rk_stage_g(const double2* __restrict__ ya, const double2* __restrict__ yb, const DevStep* __restrict__ s,
                             StageCoef sc, typename ctype<T>::type* __restrict__ out, size_t per, size_t n) {
#pragma clang fp contract(off)
    using C = typename ctype<T>::type;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        const DevStep& d = s[i / per];
        if (d.done) continue;
        const double h = d.h0;
        const double2 yv = (d.cur ? yb : ya)[i];
        double ax = 0.0, ay = 0.0;
        for (int j = 0; j < sc.nk; ++j) {
            C kv = reinterpret_cast<const C*>(sc.k[j])[i];
            ax += sc.b[j] * (double)kv.x;
            ay += sc.b[j] * (double)kv.y;
        }
        double rx = yv.x + h * ax, ry = yv.y + h * ay;
        out[i] = cmk<C>((T)rx, (T)ry);
    }
}

// y_new = y + h (c1 k1 + c3 k3 + c4 k4 + c5 k5 + c6 k6)
// err   = h (ec1 k1 + ec3 k3 + ec4 k4 + ec5 k5 + ec6 k6)
// rmax  = max_i |err_i| / (eps_abs + eps_rel |y_new_i|) over real components,
// NaN ratios skipped (GSL_MAX_DBL keeps the running max when r is NaN).
// The max is reduced with wavefront shuffles, then one atomicMax per block on
// the bit pattern of the (non-negative) double.
template <typename T>
__device__ __forceinline__ void rk_final_body(const double2* __restrict__ y, const typename ctype<T>::type* __restrict__ k1,
                                              const typename ctype<T>::type* __restrict__ k3, const typename ctype<T>::type* __restrict__ k4,
                                              const typename ctype<T>::type* __restrict__ k5, const typename ctype<T>::type* __restrict__ k6,
                                              double h, double eps_abs, double eps_rel, double2* __restrict__ ynew,
                                              unsigned long long* __restrict__ rmax_bits, size_t n,
                                              typename ctype<T>::type* __restrict__ ynew_t = nullptr) {
#pragma clang fp contract(off)
    const double c1 = 902880.0 / 7618050.0, c3 = 3953664.0 / 7618050.0, c4 = 3855735.0 / 7618050.0,
                 c5 = -1371249.0 / 7618050.0, c6 = 277020.0 / 7618050.0;
    const double e1 = 1.0 / 360.0, e3 = -128.0 / 4275.0, e4 = -2197.0 / 75240.0, e5 = 1.0 / 50.0, e6 = 2.0 / 55.0;
    double rmax = 0.0;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        double2 yv = y[i];
        auto K1 = k1[i]; auto K3 = k3[i]; auto K4 = k4[i]; auto K5 = k5[i]; auto K6 = k6[i];
        double dx = c1 * (double)K1.x + c3 * (double)K3.x + c4 * (double)K4.x + c5 * (double)K5.x + c6 * (double)K6.x;
        double dy = c1 * (double)K1.y + c3 * (double)K3.y + c4 * (double)K4.y + c5 * (double)K5.y + c6 * (double)K6.y;
        double nx = yv.x + h * dx, ny = yv.y + h * dy;
        ynew[i] = make_double2(nx, ny);
        if (ynew_t) ynew_t[i] = cmk<typename ctype<T>::type>((T)nx, (T)ny);   // the next RHS input in T
        double ex = h * (e1 * (double)K1.x + e3 * (double)K3.x + e4 * (double)K4.x + e5 * (double)K5.x + e6 * (double)K6.x);
        double ey = h * (e1 * (double)K1.y + e3 * (double)K3.y + e4 * (double)K4.y + e5 * (double)K5.y + e6 * (double)K6.y);
        double rx = fabs(ex) / fabs(eps_rel * fabs(nx) + eps_abs);
        double ry = fabs(ey) / fabs(eps_rel * fabs(ny) + eps_abs);
        if (rx > rmax) rmax = rx;
        if (ry > rmax) rmax = ry;
    }
    // wavefront (64-lane) reduction
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(rmax, off, 64);
        rmax = o > rmax ? o : rmax;
    }
    __shared__ double wmax[4];
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) wmax[wv] = rmax;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = wmax[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = wmax[w] > m ? wmax[w] : m;
        atomicMax(rmax_bits, (unsigned long long)__double_as_longlong(m));
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_rk_final(const double2* __restrict__ y, const typename ctype<T>::type* __restrict__ k1,
                                                  const typename ctype<T>::type* __restrict__ k3, const typename ctype<T>::type* __restrict__ k4,
                                                  const typename ctype<T>::type* __restrict__ k5, const typename ctype<T>::type* __restrict__ k6,
                                                  double h, double eps_abs, double eps_rel, double2* __restrict__ ynew,
                                                  unsigned long long* __restrict__ rmax_bits, size_t n,
                                                  typename ctype<T>::type* __restrict__ ynew_t) {
    rk_final_body<T>(y, k1, k3, k4, k5, k6, h, eps_abs, eps_rel, ynew, rmax_bits, n, ynew_t);
}

// the final kernel of a graph attempt, member blockIdx.y (elements
// [z per, (z + 1) per)): y from buffer cur, y_new into the other
template <typename T>
__global__ void __launch_bounds__(256) k_rk_final_g(double2* __restrict__ ya, double2* __restrict__ yb, DevStep* __restrict__ sv,
                                                    const typename ctype<T>::type* __restrict__ k1,
                                                    const typename ctype<T>::type* __restrict__ k3, const typename ctype<T>::type* __restrict__ k4,
                                                    const typename ctype<T>::type* __restrict__ k5, const typename ctype<T>::type* __restrict__ k6,
                                                    size_t per) {
    DevStep* s = sv + blockIdx.y;
    if (s->done) return;
    const size_t o = (size_t)blockIdx.y * per;
    const int cur = s->cur;
    rk_final_body<T>((cur ? yb : ya) + o, k1 + o, k3 + o, k4 + o, k5 + o, k6 + o, s->h0, s->eps_abs, s->eps_rel,
                     (cur ? ya : yb) + o, &s->rmax_bits, per);
}

// control_y and evolve_apply's accept / retry / underflow decision
static __global__ void k_step_end(DevStep* sv, int nb) {
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= nb) return;
    DevStep* s = sv + z;
    if (s->done) return;
    s->attempts++;
    double rm = __longlong_as_double((long long)s->rmax_bits);
    if (rm < 2.2250738585072014e-308) rm = 2.2250738585072014e-308;     // GSL starts rmax at DBL_MIN
    const double tn = s->final_step ? s->t1 : __dadd_rn(s->t0, s->h0);
    const double h_old = s->h0;
    double hn = h_old;
    if (hadjust_rule(rm, &hn) < 0) {
        if (fabs(hn) < fabs(h_old) && __dadd_rn(tn, hn) != tn) {
            s->rejected++;
            s->h = hn;                  // retry from t0 with the smaller step
            return;
        }
        s->h = hn;
        s->err = 1;
        s->done = 1;
        return;
    }
    s->accepted++;
    s->t = tn;
    s->h = hn;
    s->cur ^= 1;
    if (!(tn < s->t1)) s->done = 1;     // ode_solver_uv: while (t < t1)
}

// ------------------------------------------------- boundary conversions ---
// reference layout: 6 planes [Re eta][Im eta][Re u][Im u][Re v][Im v], each [ny][nx]
// half layout: 3 fields, each [kx][ky] for kx = 0..nx/2
template <typename CO>
__global__ void k_full_to_half(const double* __restrict__ full, CO* __restrict__ half, int nx, int ny) {
    __shared__ double tre[32][33], tim[32][33];
    int f = blockIdx.z;
    int kx0 = blockIdx.x * 32, ky0 = blockIdx.y * 32;
    int nxh = nx / 2 + 1;
    size_t N = (size_t)nx * ny;
    const double* re = full + (size_t)(2 * f) * N;
    const double* im = full + (size_t)(2 * f + 1) * N;
    for (int r = threadIdx.y; r < 32; r += blockDim.y) {
        int ky = ky0 + r, kx = kx0 + threadIdx.x;
        if (ky < ny && kx < nxh) {
            tre[r][threadIdx.x] = re[(size_t)ky * nx + kx];
            tim[r][threadIdx.x] = im[(size_t)ky * nx + kx];
        }
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 32; r += blockDim.y) {
        int kx = kx0 + r, ky = ky0 + threadIdx.x;
        if (ky < ny && kx < nxh) {
            CO v;
            v.x = (decltype(v.x))tre[threadIdx.x][r];
            v.y = (decltype(v.y))tim[threadIdx.x][r];
            half[(size_t)f * nxh * ny + (size_t)kx * ny + ky] = v;
        }
    }
}

// half -> full; columns kx > nx/2 are filled with the Hermitian mirror
// conj(half[(ny-ky)%ny][nx-kx]).
template <typename CI>
__global__ void k_half_to_full(const CI* __restrict__ half, double* __restrict__ full, int nx, int ny) {
    __shared__ double tre[32][33], tim[32][33];
    int f = blockIdx.z;
    int kx0 = blockIdx.x * 32, ky0 = blockIdx.y * 32;   // kx over the full width
    int nxh = nx / 2 + 1;
    size_t N = (size_t)nx * ny;
    for (int r = threadIdx.y; r < 32; r += blockDim.y) {
        int kx = kx0 + r, ky = ky0 + threadIdx.x;
        if (kx < nx && ky < ny) {
            double a, b;
            if (kx < nxh) {
                CI v = half[(size_t)f * nxh * ny + (size_t)kx * ny + ky];
                a = v.x; b = v.y;
            } else {
                int mk = (ny - ky) % ny, mx = nx - kx;
                CI v = half[(size_t)f * nxh * ny + (size_t)mx * ny + mk];
                a = v.x; b = -(double)v.y;
            }
            tre[r][threadIdx.x] = a;
            tim[r][threadIdx.x] = b;
        }
    }
    __syncthreads();
    double* re = full + (size_t)(2 * f) * N;
    double* im = full + (size_t)(2 * f + 1) * N;
    for (int r = threadIdx.y; r < 32; r += blockDim.y) {
        int ky = ky0 + r, kx = kx0 + threadIdx.x;
        if (kx < nx && ky < ny) {
            re[(size_t)ky * nx + kx] = tre[threadIdx.x][r];
            im[(size_t)ky * nx + kx] = tim[threadIdx.x][r];
        }
    }
}

template <typename CI, typename CO>
__global__ void k_convert(const CI* __restrict__ a, CO* __restrict__ b, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        CI v = a[i];
        CO o;
        o.x = (decltype(o.x))v.x;
        o.y = (decltype(o.y))v.y;
        b[i] = o;
    }
}

template <typename A, typename B>
__global__ void k_convert_scalar(const A* __restrict__ a, B* __restrict__ b, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) b[i] = (B)a[i];
}

// a fill kernel (not hipMemsetAsync) so the sequence stays graph-capturable
static __global__ void k_zero_u64(unsigned long long* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// ----------------------------------------------------------- influx ------
// natural-spline evaluation of every line signal at one time (the interval
// idx and offset dx are found on the host by the bsearch rule of GSL).
// tables are [nt][nline]: y values and half second derivatives c.
template <typename T>
__global__ void k_src_eval(const double* __restrict__ sy, const double* __restrict__ scf, const double* __restrict__ ky_,
                           const double* __restrict__ kc, int nline, int idx, double dx, double hh, T* __restrict__ osig,
                           T* __restrict__ oskew) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nline) return;
    for (int s = 0; s < 2; ++s) {
        const double* Y = s == 0 ? sy : ky_;
        const double* Cc = s == 0 ? scf : kc;
        double y0 = Y[(size_t)idx * nline + j], y1 = Y[(size_t)(idx + 1) * nline + j];
        double ci = Cc[(size_t)idx * nline + j], cip = Cc[(size_t)(idx + 1) * nline + j];
        double b = (y1 - y0) / hh - hh * (cip + 2.0 * ci) / 3.0;
        double d = (cip - ci) / (3.0 * hh);
        double v = y0 + dx * (b + dx * (ci + dx * d));
        if (s == 0) osig[j] = (T)v; else oskew[j] = (T)v;
    }
}

// k_src_eval at the stage time t0 + frac h0 of a graph attempt: the interval
// is found on the device by the same bisection (gsl_interp_bsearch)
template <typename T>
__global__ void k_src_eval_g(const DevStep* __restrict__ s, double frac, const double* __restrict__ tsig, int nt,
                             const double* __restrict__ sy, const double* __restrict__ scf, const double* __restrict__ ky_,
                             const double* __restrict__ kc, int nline, T* __restrict__ osig, T* __restrict__ oskew) {
    if (s->done) return;
    const double t = __dadd_rn(s->t0, __dmul_rn(frac, s->h0));
    int ilo = 0, ihi = nt - 1;
    while (ihi > ilo + 1) {
        int i = (ihi + ilo) / 2;
        if (tsig[i] > t) ihi = i; else ilo = i;
    }
    const double hh = tsig[ilo + 1] - tsig[ilo], dx = t - tsig[ilo];
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nline) return;
    for (int q = 0; q < 2; ++q) {
        const double* Y = q == 0 ? sy : ky_;
        const double* Cc = q == 0 ? scf : kc;
        double y0 = Y[(size_t)ilo * nline + j], y1 = Y[(size_t)(ilo + 1) * nline + j];
        double ci = Cc[(size_t)ilo * nline + j], cip = Cc[(size_t)(ilo + 1) * nline + j];
        double b = (y1 - y0) / hh - hh * (cip + 2.0 * ci) / 3.0;
        double d = (cip - ci) / (3.0 * hh);
        double v = y0 + dx * (b + dx * (ci + dx * d));
        if (q == 0) osig[j] = (T)v; else oskew[j] = (T)v;
    }
}

// ------------------------------------------------ stand-alone transforms --
// forward 2-D real transform: physical [ny][nx] -> half spectrum [kx][ky]
template <typename T, int NX, int NY>
__global__ void HW_XK_BOUNDS k_t_xfwd(const T* __restrict__ phys, typename ctype<T>::type* X,
                                                            const typename ctype<T>::type* twx) {
    HW_XK_PROLOGUE_TW(twx)
    C f[G::PP];
    ph_load<T, NX, NY>(f, phys, y0, tid);
    x_fwd_store<T, NX, NY>(f, X, y0, lds, twxs, tid);
}

template <typename T, int NX, int NY>
__global__ void __launch_bounds__((YGeo<T, NY>::NT)) k_t_yfwd(const typename ctype<T>::type* X, typename ctype<T>::type* S,
                                                            const typename ctype<T>::type* twy) {
    HW_YK_PROLOGUE_TW(twy)
    C v[G::PE];
    y_load<T, NX, NY>(v, X, nullptr, col0, tid);
    y_fwd_regs<T, NX, NY>(v, lds, twys, tid);
    y_store<T, NX, NY>(S, v, col0, tid);
}

// inverse: half spectrum (reference projection semantics) -> physical
template <typename T, int NX, int NY>
__global__ void __launch_bounds__((YGeo<T, NY>::NT)) k_t_yinv(const typename ctype<T>::type* S, typename ctype<T>::type* X,
                                                            const typename ctype<T>::type* twy) {
    HW_YK_PROLOGUE_TW(twy)
    C v[G::PE];
    y_load<T, NX, NY>(v, S, nullptr, col0, tid);
    y_inv_regs<T, NX, NY>(v, lds, twys, col0, tid);
    y_store<T, NX, NY>(X, v, col0, tid);
}

template <typename T, int NX, int NY>
__global__ void HW_XK_BOUNDS k_t_xinv(const typename ctype<T>::type* X, T* phys, T* nyq_out,
                                                            const typename ctype<T>::type* twx, const int* gate = nullptr) {
    if (gate && *gate == 0) return;
    HW_XK_PROLOGUE_TW(twx)
    C f[G::PP];
    x_inverse<T, NX, NY>(X, y0, lds, twxs, f, sc, nyq, tid);
    ph_store<T, NX, NY>(phys, f, y0, tid);
    if (nyq_out && tid < G::RR) nyq_out[y0 + tid] = nyq[tid];
}

}  // namespace hw
