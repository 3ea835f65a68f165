// kin1d.hip -- internal flow kinematics of the 1-D model at an output time:
// kinematic_modul (1d/kinematic_modul.c:1-247) behind hw1d_kinematics.
//
// Per node j (k_kin_nodes): H = eta + depth, the two-depth weights
// gamma_-(H), gamma_+(H) from natural splines over the sinhcosh2IP table,
// cosh(kH) and tanh(kH) as their weighted sums at the table's end depths,
// phi^ = -i u^ / k and the one-sided / central gradients of H and depth
// (1d/kinematic_modul.c:97-117, gradient at 1d/functions.c:1850-1856).  As in
// the reference, the node index j also indexes the wavenumber k[j]: the
// depth-dependent ratios multiply the spectrum entry of the same index.
//
// Per output level z (k_kin_level, one block each): cosh / sinh(k DZ) with
// DZ = depth + z clipped at 0 (flat bottom, or one DZ over the whole level:
// cosh / sinh of k (z + depth[0])), then the four spectra of
// 1d/kinematic_modul.c:155-166 and their inverse transforms
// (ifft_complex_real_nx, 1d/functions.c:1642-1661: real part of the backward
// DFT / n).  Two real outputs come from one complex transform: Re F^-1 X =
// F^-1 Herm(X), Herm(X)[k] = (X[k] + conj X[-k]) / 2, so one LDS Stockham
// transform of Herm(A) + i Herm(B) gives both.
//
// Per output point (k_kin_out): the mask above the surface and below the
// bottom (NaN), dtPhi, dxPhi, dzPhi, the pressure exactly as the reference
// writes it (-dtPhi - g z - (dxPhi^2 + 2 dzPhi) / 2, 1d/kinematic_modul.c:199)
// and the accelerations (v - v_prev) / (t - tprev) (:212-217).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "../../include/hawassi_hip.h"
#include "hw_fft.h"

namespace hw {

void set_error(const char* fmt, ...);

namespace {

struct KinArgs {
    int n, nz, ix0, nxout, flat, ntab;
    double hmin, hplus, grav, dx, t, tprev;
    const double *depth, *k, *z;
    const double *htab, *ghm, *ghm_c, *ghp, *ghp_c;        // H table, gamma_-, gamma_+ and spline coefficients
    const double *dztab, *gdm, *gdm_c, *gdp, *gdp_c;       // DZ table
    const double *eta, *deta, *velx, *velz;
    const double2 *uhat, *dphihat;
    const double2* tw;                                     // exp(-2 pi i m / n)
    double *ckh, *tkh, *gradh, *gradd;                     // per node
    double2* phihat;
    double* scr;                                           // [nz][4][n]: factor, dtPhi, dzPhi, dxPhi before the mask
    double* out;                                           // [6][nxout][nz]
};

// gsl_spline_eval of a natural cubic spline (gsl_interp_cspline): interval
// by bisection, y_i + dx (b + dx (c_i + dx d)) with c the half second
// derivatives
__device__ double kin_spline(const double* x, const double* y, const double* c, int n, double t) {
    int lo = 0, hi = n - 1;
    while (hi > lo + 1) {
        int m = (hi + lo) >> 1;
        if (x[m] > t) hi = m;
        else lo = m;
    }
    const double h = x[lo + 1] - x[lo], dx = t - x[lo], dy = y[lo + 1] - y[lo];
    const double ci = c[lo], cip = c[lo + 1];
    const double b = dy / h - h * (cip + 2.0 * ci) / 3.0, d = (cip - ci) / (3.0 * h);
    return y[lo] + dx * (b + dx * (ci + dx * d));
}

__global__ void k_kin_nodes(KinArgs a) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.n) return;
    const int n = a.n;
    auto Hat = [&](int i) { return a.eta[i] + a.depth[i]; };
    const double H = Hat(j), kk = a.k[j];
    const double gm = kin_spline(a.htab, a.ghm, a.ghm_c, a.ntab, H);
    const double gp = kin_spline(a.htab, a.ghp, a.ghp_c, a.ntab, H);
    a.ckh[j] = gm * cosh(kk * a.hmin) + gp * cosh(kk * a.hplus);
    a.tkh[j] = gm * tanh(kk * a.hmin) + gp * tanh(kk * a.hplus);
    const double2 u = a.uhat[j];
    a.phihat[j] = kk == 0.0 ? make_double2(0.0, 0.0) : make_double2(u.y / kk, -u.x / kk);
    if (j == 0) {
        a.gradh[j] = (Hat(1) - Hat(0)) / a.dx;
        a.gradd[j] = (a.depth[1] - a.depth[0]) / a.dx;
    } else if (j == n - 1) {
        a.gradh[j] = (Hat(n - 1) - Hat(n - 2)) / a.dx;
        a.gradd[j] = (a.depth[n - 1] - a.depth[n - 2]) / a.dx;
    } else {
        a.gradh[j] = (Hat(j + 1) - Hat(j - 1)) / (2 * a.dx);
        a.gradd[j] = (a.depth[j + 1] - a.depth[j - 1]) / (2 * a.dx);
    }
}

// the pair p of spectra at index j of level z: p = 0 (factor, dtPhi),
// p = 1 (dzPhi, dxPhi)  (1d/kinematic_modul.c:155-166)
__device__ void kin_spectra(const KinArgs& a, int j, double z, bool one_dz, double dzmin, double dzplus, int p,
                            double2& s0, double2& s1) {
    const double kk = a.k[j];
    double ckdz, skdz;
    if (one_dz) {
        const double arg = kk * (z + a.depth[0]);
        ckdz = cosh(arg);
        skdz = sinh(arg);
    } else {
        double d = a.depth[j] + z;
        if (d < 0) d = 0;
        double gm = kin_spline(a.dztab, a.gdm, a.gdm_c, a.ntab, d);
        double gp = kin_spline(a.dztab, a.gdp, a.gdp_c, a.ntab, d);
        if (d == dzmin) { gm = 1; gp = 0; }
        if (d == dzplus) { gm = 0; gp = 1; }
        ckdz = gm * cosh(kk * dzmin) + gp * cosh(kk * dzplus);
        skdz = gm * sinh(kk * dzmin) + gp * sinh(kk * dzplus);
    }
    const double ckh = a.ckh[j];
    const double ratio = ckdz / ckh;
    const double2 ph = a.phihat[j];
    if (p == 0) {
        const double f = -kk * a.tkh[j];
        s0 = make_double2(ph.x * ratio * f, ph.y * ratio * f);
        const double2 dp = a.dphihat[j];
        s1 = make_double2(dp.x * ratio, dp.y * ratio);
    } else {
        s0 = make_double2(ph.x * kk * skdz / ckh, ph.y * kk * skdz / ckh);
        s1 = make_double2(-kk * ph.y * ratio, kk * ph.x * ratio);
    }
}

template <int N, int NT> __global__ void __launch_bounds__(NT) k_kin_level(KinArgs a) {
    extern __shared__ double2 kin_lds[];
    __shared__ double red[2][NT / 64];
    const int zz = blockIdx.x, tid = threadIdx.x;
    const double z = a.z[zz];
    double dzmin = 0.0, dzplus = 0.0;
    bool one_dz = a.flat != 0;
    if (!one_dz) {
        // DZplus / DZmin of the level (1d/kinematic_modul.c:131-138)
        double mn = INFINITY, mx = -INFINITY;
        for (int j = tid; j < N; j += NT) {
            double d = a.depth[j] + z;
            if (d < 0) d = 0;
            mn = fmin(mn, d);
            mx = fmax(mx, d);
        }
        for (int o = 32; o > 0; o >>= 1) {
            mn = fmin(mn, __shfl_xor(mn, o));
            mx = fmax(mx, __shfl_xor(mx, o));
        }
        if ((tid & 63) == 0) { red[0][tid >> 6] = mn; red[1][tid >> 6] = mx; }
        __syncthreads();
        mn = red[0][0];
        mx = red[1][0];
        for (int w = 1; w < NT / 64; ++w) { mn = fmin(mn, red[0][w]); mx = fmax(mx, red[1][w]); }
        dzmin = mn < 0 ? 0.0 : mn;
        dzplus = mx;
        one_dz = dzmin == dzplus;
    }
    for (int p = 0; p < 2; ++p) {
        // one thread per index pair (j, N - j): Herm(X)[N - j] is the
        // conjugate of Herm(X)[j], so each spectrum entry is evaluated once
        for (int j = tid; j <= N / 2; j += NT) {
            const int jm = j == 0 ? 0 : N - j;
            double2 aj, bj, am, bm;
            kin_spectra(a, j, z, one_dz, dzmin, dzplus, p, aj, bj);
            if (jm != j) kin_spectra(a, jm, z, one_dz, dzmin, dzplus, p, am, bm);
            else { am = aj; bm = bj; }
            const double hx = 0.5 * (aj.x + am.x), hy = 0.5 * (aj.y - am.y);
            const double gx = 0.5 * (bj.x + bm.x), gy = 0.5 * (bj.y - bm.y);
            kin_lds[j] = make_double2(hx - gy, hy + gx);
            if (jm != j) kin_lds[jm] = make_double2(hx + gy, gx - hy);
        }
        __syncthreads();
        fft_lds<double2, N, 1, NT, 1>(kin_lds, N, a.tw, 1, tid);   // backward, ends with a barrier
        double* o0 = a.scr + ((size_t)zz * 4 + 2 * p) * N;
        double* o1 = o0 + N;
        for (int j = tid; j < N; j += NT) {
            const double2 v = kin_lds[j];
            o0[j] = v.x / N;
            o1[j] = v.y / N;
        }
        __syncthreads();
    }
}

__global__ void k_kin_out(KinArgs a) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = a.nxout * a.nz;
    if (idx >= m) return;
    const int xo = idx / a.nz, zz = idx % a.nz, j = a.ix0 + xo, n = a.n;
    const double* s = a.scr + (size_t)zz * 4 * n;
    const double fac = s[j], dti = s[n + j], dzi = s[2 * n + j], dxi = s[3 * n + j];
    const double z = a.z[zz];
    double dt, dxv, dzv;
    if (z > a.eta[j] || z < -a.depth[j]) {
        dt = dxv = dzv = __builtin_nan("");
    } else {
        dt = dti + a.deta[j] * fac;
        dxv = dxi + a.gradh[j] * fac + a.gradd[j] * dzi;
        dzv = dzi;
    }
    const double P = -dt - a.grav * z - 0.5 * (dxv * dxv + 2 * dzv);
    const double vx0 = a.velx ? a.velx[idx] : 0.0, vz0 = a.velz ? a.velz[idx] : 0.0;
    a.out[idx] = dt;
    a.out[m + idx] = dxv;
    a.out[2 * m + idx] = dzv;
    a.out[3 * m + idx] = P;
    a.out[4 * m + idx] = (dxv - vx0) / (a.t - a.tprev);
    a.out[5 * m + idx] = (dzv - vz0) / (a.t - a.tprev);
}

constexpr int KIN_NT = 256;

template <int N> hipError_t kin_level_launch(const KinArgs& a, hipStream_t s) {
    const size_t lds = sizeof(double2) * N;
    hipError_t e = hipFuncSetAttribute((const void*)k_kin_level<N, KIN_NT>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    k_kin_level<N, KIN_NT><<<a.nz, KIN_NT, lds, s>>>(a);
    return hipGetLastError();
}

hipError_t kin_level(const KinArgs& a, hipStream_t s) {
    switch (a.n) {
        case 256: return kin_level_launch<256>(a, s);
        case 512: return kin_level_launch<512>(a, s);
        case 1024: return kin_level_launch<1024>(a, s);
        case 2000: return kin_level_launch<2000>(a, s);
        case 2048: return kin_level_launch<2048>(a, s);
        case 4096: return kin_level_launch<4096>(a, s);
        case 8192: return kin_level_launch<8192>(a, s);
        default: return hipErrorInvalidValue;
    }
}

// natural cubic spline half second derivatives (gsl_interp_cspline: c_0 =
// c_{n-1} = 0), tridiagonal system by the Thomas algorithm
void spline_c(const double* x, const double* y, int n, double* c) {
    for (int i = 0; i < n; ++i) c[i] = 0.0;
    if (n < 3) return;
    const int m = n - 2;
    std::vector<double> h(n - 1), diag(m), rhs(m), off(m);
    for (int i = 0; i < n - 1; ++i) h[i] = x[i + 1] - x[i];
    for (int i = 0; i < m; ++i) {
        diag[i] = 2.0 * (h[i] + h[i + 1]);
        off[i] = h[i + 1];
        rhs[i] = 3.0 * ((y[i + 2] - y[i + 1]) / h[i + 1] - (y[i + 1] - y[i]) / h[i]);
    }
    for (int i = 1; i < m; ++i) {
        const double w = off[i - 1] / diag[i - 1];
        diag[i] -= w * off[i - 1];
        rhs[i] -= w * rhs[i - 1];
    }
    c[m] = rhs[m - 1] / diag[m - 1];
    for (int i = m - 2; i >= 0; --i) c[i + 1] = (rhs[i] - off[i] * c[i + 2]) / diag[i];
}

struct DevAlloc {
    void* p = nullptr;
    ~DevAlloc() { if (p) (void)hipFree(p); }
};

}  // namespace

}  // namespace hw

using hw::set_error;

#define KIN_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
            if (stream) (void)hipStreamDestroy(stream);                                   \
            return HW_ERR_HIP;                                                            \
        }                                                                                 \
    } while (0)

extern "C" int hw1d_kinematics(const hw_kin1d_config* c, const double* eta, const double* u_hat, const double* deta,
                               const double* dphi_hat, double t, double tprev, const double* velx, const double* velz,
                               double* out) {
    if (!c || !eta || !u_hat || !deta || !dphi_hat || !out) {
        set_error("hw1d_kinematics: config, eta, u_hat, deta, dphi_hat and out are required");
        return HW_ERR_ARG;
    }
    const int n = c->n;
    if (!(n == 256 || n == 512 || n == 1024 || n == 2000 || n == 2048 || n == 4096 || n == 8192)) {
        set_error("hw1d_kinematics: n = %d (supported: powers of two 256..8192 and 2000)", n);
        return HW_ERR_ARG;
    }
    if (c->ix0 < 0 || c->nxout < 1 || c->ix0 + c->nxout > n || c->nz < 1 || c->ntab < 3 || !(c->dx > 0)) {
        set_error("hw1d_kinematics: window %d + %d of %d nodes, %d levels, %d table points, dx %g", c->ix0,
                  c->nxout, n, c->nz, c->ntab, c->dx);
        return HW_ERR_ARG;
    }
    if (!c->depth || !c->k || !c->z || !c->htab || !c->gam_hmin || !c->gam_hplus || !c->dztab || !c->gam_dzmin ||
        !c->gam_dzplus) {
        set_error("hw1d_kinematics: depth, k, z and both gamma tables are required");
        return HW_ERR_ARG;
    }
    for (int i = 0; i + 1 < c->ntab; ++i)
        if (!(c->htab[i + 1] > c->htab[i]) || !(c->dztab[i + 1] > c->dztab[i])) {
            set_error("hw1d_kinematics: spline tables must increase strictly");
            return HW_ERR_ARG;
        }
    // gsl_spline_eval stops the reference with a domain error outside the table
    const double hlo = c->htab[0], hhi = c->htab[c->ntab - 1];
    for (int j = 0; j < n; ++j) {
        const double H = eta[j] + c->depth[j];
        if (!(H >= hlo && H <= hhi)) {
            set_error("hw1d_kinematics: eta + depth = %g at node %d is outside the table [%g, %g]", H, j, hlo, hhi);
            return HW_ERR_ARG;
        }
    }
    int ndev = 0;
    hipError_t ge = hipGetDeviceCount(&ndev);
    if (ge != hipSuccess || ndev < 1) {
        set_error("no HIP device available (%s)", ge == hipSuccess ? "count 0" : hipGetErrorString(ge));
        return HW_ERR_NOGPU;
    }
    const int nt = c->ntab, nz = c->nz;
    const size_t m = (size_t)c->nxout * nz;
    // host staging: tables with their spline coefficients, the twiddles
    std::vector<double> tabs(10 * (size_t)nt);
    double* h = tabs.data();
    for (int i = 0; i < nt; ++i) {
        h[i] = c->htab[i]; h[nt + i] = c->gam_hmin[i]; h[3 * nt + i] = c->gam_hplus[i];
        h[5 * nt + i] = c->dztab[i]; h[6 * nt + i] = c->gam_dzmin[i]; h[8 * nt + i] = c->gam_dzplus[i];
    }
    hw::spline_c(h, h + nt, nt, h + 2 * nt);
    hw::spline_c(h, h + 3 * nt, nt, h + 4 * nt);
    hw::spline_c(h + 5 * nt, h + 6 * nt, nt, h + 7 * nt);
    hw::spline_c(h + 5 * nt, h + 8 * nt, nt, h + 9 * nt);
    std::vector<double2> tw(n);
    for (int q = 0; q < n; ++q) {
        const double a = -2.0 * M_PI * (double)q / (double)n;
        tw[q] = make_double2(std::cos(a), std::sin(a));
    }
    // one device block: [tables][depth k z eta deta][uhat dphihat tw][ckh tkh gradh gradd][phihat][velx velz][scr][out]
    size_t off = 0;
    auto take = [&](size_t count) { size_t o = off; off += count; return o; };
    const size_t o_tab = take(10 * (size_t)nt), o_dep = take(n), o_k = take(n), o_z = take(nz), o_eta = take(n),
                 o_deta = take(n), o_uh = take(2 * (size_t)n), o_dph = take(2 * (size_t)n), o_tw = take(2 * (size_t)n),
                 o_ckh = take(n), o_tkh = take(n), o_gh = take(n), o_gd = take(n), o_ph = take(2 * (size_t)n),
                 o_vx = take(m), o_vz = take(m), o_scr = take((size_t)nz * 4 * n), o_out = take(6 * m);
    hipStream_t stream = nullptr;
    hw::DevAlloc buf;
    KIN_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    KIN_TRY(hipMalloc(&buf.p, sizeof(double) * off));
    double* d = static_cast<double*>(buf.p);
    auto up = [&](size_t o, const void* src, size_t count) {
        return hipMemcpyAsync(d + o, src, sizeof(double) * count, hipMemcpyHostToDevice, stream);
    };
    KIN_TRY(up(o_tab, tabs.data(), 10 * (size_t)nt));
    KIN_TRY(up(o_dep, c->depth, n));
    KIN_TRY(up(o_k, c->k, n));
    KIN_TRY(up(o_z, c->z, nz));
    KIN_TRY(up(o_eta, eta, n));
    KIN_TRY(up(o_deta, deta, n));
    KIN_TRY(up(o_uh, u_hat, 2 * (size_t)n));
    KIN_TRY(up(o_dph, dphi_hat, 2 * (size_t)n));
    KIN_TRY(up(o_tw, tw.data(), 2 * (size_t)n));
    if (velx) KIN_TRY(up(o_vx, velx, m));
    if (velz) KIN_TRY(up(o_vz, velz, m));
    hw::KinArgs a{};
    a.n = n; a.nz = nz; a.ix0 = c->ix0; a.nxout = c->nxout; a.flat = c->flat != 0; a.ntab = nt;
    a.hmin = c->hmin; a.hplus = c->hplus; a.grav = c->grav; a.dx = c->dx; a.t = t; a.tprev = tprev;
    a.depth = d + o_dep; a.k = d + o_k; a.z = d + o_z;
    a.htab = d + o_tab; a.ghm = d + o_tab + nt; a.ghm_c = d + o_tab + 2 * nt; a.ghp = d + o_tab + 3 * nt;
    a.ghp_c = d + o_tab + 4 * nt; a.dztab = d + o_tab + 5 * nt; a.gdm = d + o_tab + 6 * nt;
    a.gdm_c = d + o_tab + 7 * nt; a.gdp = d + o_tab + 8 * nt; a.gdp_c = d + o_tab + 9 * nt;
    a.eta = d + o_eta; a.deta = d + o_deta;
    a.velx = velx ? d + o_vx : nullptr; a.velz = velz ? d + o_vz : nullptr;
    a.uhat = reinterpret_cast<const double2*>(d + o_uh);
    a.dphihat = reinterpret_cast<const double2*>(d + o_dph);
    a.tw = reinterpret_cast<const double2*>(d + o_tw);
    a.ckh = d + o_ckh; a.tkh = d + o_tkh; a.gradh = d + o_gh; a.gradd = d + o_gd;
    a.phihat = reinterpret_cast<double2*>(d + o_ph);
    a.scr = d + o_scr; a.out = d + o_out;
    hw::k_kin_nodes<<<(n + 255) / 256, 256, 0, stream>>>(a);
    KIN_TRY(hipGetLastError());
    KIN_TRY(hw::kin_level(a, stream));
    hw::k_kin_out<<<(unsigned)((m + 255) / 256), 256, 0, stream>>>(a);
    KIN_TRY(hipGetLastError());
    KIN_TRY(hipMemcpyAsync(out, d + o_out, sizeof(double) * 6 * m, hipMemcpyDeviceToHost, stream));
    KIN_TRY(hipStreamSynchronize(stream));
    hipError_t de = hipStreamDestroy(stream);
    stream = nullptr;
    KIN_TRY(de);
    return HW_OK;
}
