// breaking2d.h -- kinematic breaking criterion, crest tracking and the
// eddy-viscosity source of the 2-D model (reference: 2d/breaking.c and
// fun_sbreaking, 2d/fun_odesolver.c:985-1019).
//
// Work split: every full-grid scan runs on the GPU (peak search, Hilbert
// local wavenumber, crest re-location, break-node marking, flux and its
// clipped gradient, transforms); the crest bookkeeping of breaking_process
// (a few scalars per crest: the flagbr state machine, new-crest tests) runs
// on the host between kernels.
//
// Break nodes: the reference visits crests in order and a node takes B from
// the first crest that gives it B > 0 (B == 0 test, 2d/breaking.c:476).  The
// device gives every crest one block; candidates do atomicMin(owner, crest)
// and a second pass lets only the owning crest write B, which is the same
// result independent of scheduling.
#pragma once
#include "rhs2d_kernels.h"

namespace hw {

struct BrkPeak {
    int x, y;
};


// a peak that passed the crest test, appended by k_brk_crit in any order
// (the host restores the reference's peak order)
struct BrkCand {
    int x, y, prevnode, kwadran;
    double U, theta;
};

// per-crest input/output of the tracking kernels
struct BrkCrest {
    int xNow, yNow;
    double theta;
    int kwadran;
    double Ustar, UF;
    int mode;          // 1: new crest (no re-location), 2: re-locate
    // outputs
    int found, nb;
};

struct BrkGeom {
    const double* x;   // dominfo.x
    const double* y;   // dominfo.y
    double R;          // search radius
    int nx, ny;
};

__device__ __forceinline__ int kwadran_of(double ux, double uy) {
    if (ux >= 0 && uy >= 0) return 1;
    if (ux < 0 && uy >= 0) return 2;
    if (ux < 0 && uy < 0) return 3;
    return 4;
}

// funBr_Char (2d/breaking.c:359-440): disk of radius R around (xp, yp),
// cut by the quadrant sector of spread 80 degrees
__device__ __forceinline__ bool crest_char(const BrkGeom& g, int i, int j, double xp, double yp, double th, int kw) {
    const double sp = 80.0 * M_PI / 180.0;
    double al, be;
    if (kw == 1 || kw == 3) { al = th - sp; be = (sp + th) - M_PI / 2.0; }
    else { al = sp - th; be = (M_PI / 2.0 - th) - sp; }
    double xc = g.x[j] - xp, yc = g.y[i] - yp;
    if (!((xc * xc + yc * yc) <= g.R * g.R)) return false;
    double XX = xc * cos(al) + yc * sin(al);
    double YY = yc * cos(be) - xc * sin(be);
    if (kw == 1 && (XX < 0 || YY < 0)) return false;
    if (kw == 2 && (XX > 0 || YY < 0)) return false;
    if (kw == 3 && (XX > 0 || YY > 0)) return false;
    if (kw == 4 && (XX < 0 || YY > 0)) return false;
    return true;
}

// fun_peaks (2d/breaking.c:55-133): 3x3 maxima within 1e-3 above MaxEtaInit
template <typename T>
__global__ void k_brk_peaks(const T* __restrict__ eta, const T* __restrict__ pbc, double thr, int nx, int ny,
                            BrkPeak* __restrict__ out, int* __restrict__ count, int maxpk) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int yy = blockIdx.y;
    if (x >= nx) return;
    auto E = [&](int i, int j) -> double { return (double)eta[(size_t)i * nx + j] * (double)pbc[(size_t)i * nx + j]; };
    double c = E(yy, x);
    if (!(c > thr)) return;
    int i0, i1, j0, j1, need;
    bool xin = x >= 1 && x <= nx - 2, yin = yy >= 1 && yy <= ny - 2;
    if (xin && yin) { i0 = -1; i1 = 1; j0 = -1; j1 = 1; need = 9; }
    else if (x == 0 && yin) { i0 = -1; i1 = 1; j0 = 0; j1 = 1; need = 6; }
    else if (x == nx - 1 && yin) { i0 = -1; i1 = 1; j0 = -1; j1 = 0; need = 6; }
    else if (yy == 0 && xin) { i0 = 0; i1 = 1; j0 = -1; j1 = 1; need = 6; }
    else if (yy == ny - 1 && xin) { i0 = -1; i1 = 0; j0 = -1; j1 = 1; need = 6; }
    else return;   // corners are never tested
    int cnt = 0;
    for (int i = i0; i <= i1; ++i)
        for (int j = j0; j <= j1; ++j) cnt += (E(yy + i, x + j) - c <= 0.001) ? 1 : 0;
    if (cnt != need) return;
    int k = atomicAdd(count, 1);
    if (k < maxpk) { out[k].x = x; out[k].y = yy; }
}

// Hilbert transform input: -i sign(kx + ky) eta^, inverse y with projection
template <typename T, int NX, int NY>
__global__ void __launch_bounds__((YGeo<T, NY>::NT)) k_brk_hilb_y(Rhs2DArgs<T> a, typename ctype<T>::type* X,
                                                                  const int* __restrict__ gate) {
    if (*gate == 0) return;   // no peak: the Hilbert transform is not needed
    HW_YK_PROLOGUE
    C v[G::PE];
    y_load<T, NX, NY>(v, a.eta_hat, nullptr, col0, tid);
#pragma unroll
    for (int e = 0; e < G::PE; ++e) {
        int kx, ky; size_t gi;
        YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
        if (kx < NX / 2 + 1) {
            T s = (a.kxh[kx] + a.ky[ky]) < 0 ? (T)-1 : (T)1;
            v[e] = cmul_i<-1>(cscale(v[e], s));
        }
    }
    y_inv_regs<T, NX, NY>(v, lds, twys, col0, tid);
    y_store<T, NX, NY>(X, v, col0, tid);
}

// breaking_criterion (2d/breaking.c:253-305) at the peaks
// cnt[0] = number of peaks (k_brk_peaks), cnt[1] = candidates appended here
template <typename T>
__global__ void k_brk_crit(const BrkPeak* __restrict__ pk, int* __restrict__ cnt, int maxpk, const T* __restrict__ eta,
                           const T* __restrict__ H, const T* __restrict__ u, const T* __restrict__ v,
                           const T* __restrict__ depth, const T* __restrict__ Bprev, double dx, double dy, double U0,
                           double KBC, int nx, int ny, BrkCand* __restrict__ out) {
    const int npk = min(cnt[0], maxpk);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < npk; k += gridDim.x * blockDim.x) {
    int ix = pk[k].x, iy = pk[k].y;
    auto at = [&](const T* A, int i, int j) -> double { return (double)A[(size_t)i * nx + j]; };
    // gradient2D (2d/functions1.c:868-885)
    auto gx = [&](const T* A) -> double {
        if (ix == 0) return (at(A, iy, 1) - at(A, iy, 0)) / dx;
        if (ix == nx - 1) return (at(A, iy, nx - 1) - at(A, iy, nx - 2)) / dx;
        return (at(A, iy, ix + 1) - at(A, iy, ix - 1)) / (2 * dx);
    };
    auto gy = [&](const T* A) -> double {
        if (iy == 0) return (at(A, 1, ix) - at(A, 0, ix)) / dy;
        if (iy == ny - 1) return (at(A, ny - 1, ix) - at(A, ny - 2, ix)) / dy;
        return (at(A, iy + 1, ix) - at(A, iy - 1, ix)) / (2 * dy);
    };
    double e = at(eta, iy, ix), h = at(H, iy, ix);
    double Kx = (e * gx(H) - h * gx(eta)) / (e * e + h * h);
    double Ky = (e * gy(H) - h * gy(eta)) / (e * e + h * h);
    double Kk = sqrt(Kx * Kx + Ky * Ky);
    double D = at(depth, iy, ix) + e;
    const double g = 9.80665;
    double Cc = Kk == 0 ? sqrt(g * D) + U0 : (sqrt(g * Kk * tanh(Kk * D)) + Kk * U0) / Kk;
    double uu = at(u, iy, ix), vv = at(v, iy, ix);
    double U = sqrt(uu * uu + vv * vv);
    if (fabs(U) >= KBC * fabs(Cc)) {
        BrkCand r;
        r.x = ix;
        r.y = iy;
        r.prevnode = at(Bprev, iy, ix) > 0 ? 1 : 0;
        r.U = fabs(U);
        r.theta = fabs(atan(vv / uu));
        r.kwadran = kwadran_of(uu, vv);
        int slot = atomicAdd(cnt + 1, 1);
        if (slot < maxpk) out[slot] = r;
    }
    }
}

// one block per crest: re-locate (mode 2) by the first row-major maximum of
// eta * Char * pb.Char, then mark break-node candidates with atomicMin(owner)
template <typename T>
__global__ void __launch_bounds__(256) k_brk_track(BrkCrest* __restrict__ cr, int ncr, BrkGeom g, const T* __restrict__ eta,
                                                   const T* __restrict__ u, const T* __restrict__ v,
                                                   const T* __restrict__ pbc, int* __restrict__ owner) {
    int lx = blockIdx.x;
    if (lx >= ncr) return;
    __shared__ double smax[256];
    __shared__ int sidx[256];
    __shared__ int s_x, s_y, s_found;
    __shared__ double s_th;
    __shared__ int s_kw;
    BrkCrest c = cr[lx];
    const int nx = g.nx, ny = g.ny;
    double dxg = g.x[1] - g.x[0], dyg = g.y[1] - g.y[0];
    int rx = (int)(g.R / dxg) + 2, ry = (int)(g.R / dyg) + 2;
    if (threadIdx.x == 0) { s_x = c.xNow; s_y = c.yNow; s_th = c.theta; s_kw = c.kwadran; s_found = 1; }
    __syncthreads();
    if (c.mode == 2) {
        double xp = g.x[c.xNow], yp = g.y[c.yNow];
        int i0 = max(0, c.yNow - ry), i1 = min(ny - 1, c.yNow + ry);
        int j0 = max(0, c.xNow - rx), j1 = min(nx - 1, c.xNow + rx);
        int w = j1 - j0 + 1, cnt = w * (i1 - i0 + 1);
        double best = -1e300;
        int bidx = 0x7fffffff;
        for (int p = threadIdx.x; p < cnt; p += blockDim.x) {
            int i = i0 + p / w, j = j0 + p % w;
            size_t q = (size_t)i * nx + j;
            double val = crest_char(g, i, j, xp, yp, c.theta, c.kwadran) ? (double)eta[q] * (double)pbc[q] : 0.0;
            int lin = i * nx + j;
            if (val > best || (val == best && lin < bidx)) { best = val; bidx = lin; }
        }
        smax[threadIdx.x] = best;
        sidx[threadIdx.x] = bidx;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) {
                double ob = smax[threadIdx.x + s];
                int oi = sidx[threadIdx.x + s];
                if (ob > smax[threadIdx.x] || (ob == smax[threadIdx.x] && oi < sidx[threadIdx.x])) {
                    smax[threadIdx.x] = ob;
                    sidx[threadIdx.x] = oi;
                }
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            // idx_max_matrix starts from element (0,0) and needs a strict increase
            double t0 = crest_char(g, 0, 0, xp, yp, c.theta, c.kwadran) ? (double)eta[0] * (double)pbc[0] : 0.0;
            if (!(smax[0] > t0)) {
                s_found = 0;
            } else {
                s_x = sidx[0] % nx;
                s_y = sidx[0] / nx;
                size_t q = (size_t)s_y * nx + s_x;
                double uu = u[q], vv = v[q];
                s_th = fabs(atan(vv / uu));
                s_kw = kwadran_of(uu, vv);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        cr[lx].found = s_found;
        cr[lx].xNow = s_x; cr[lx].yNow = s_y; cr[lx].theta = s_th; cr[lx].kwadran = s_kw;
    }
    if (!s_found) return;
    // funBr_BreakNodes_char (2d/breaking.c:456-503): candidates
    double xp = g.x[s_x], yp = g.y[s_y];
    int i0 = max(0, s_y - ry), i1 = min(ny - 1, s_y + ry);
    int j0 = max(0, s_x - rx), j1 = min(nx - 1, s_x + rx);
    int w = j1 - j0 + 1, cnt = w * (i1 - i0 + 1);
    for (int p = threadIdx.x; p < cnt; p += blockDim.x) {
        int i = i0 + p / w, j = j0 + p % w;
        size_t q = (size_t)i * nx + j;
        if (!crest_char(g, i, j, xp, yp, s_th, s_kw)) continue;
        double uu = u[q], vv = v[q], U = sqrt(uu * uu + vv * vv);
        if (U < c.UF || (double)eta[q] < 0) continue;
        if ((double)pbc[q] != 1.0) continue;
        double b = U >= 2 * c.Ustar ? 1.0 : (U <= c.Ustar ? 0.0 : U / c.Ustar - 1.0);
        if (b > 0) atomicMin(owner + q, lx);
    }
}

// second pass: the owning crest writes B and counts its break nodes
template <typename T>
__global__ void __launch_bounds__(256) k_brk_resolve(BrkCrest* __restrict__ cr, int ncr, BrkGeom g, const T* __restrict__ eta,
                                                     const T* __restrict__ u, const T* __restrict__ v,
                                                     const T* __restrict__ pbc, const int* __restrict__ owner, T* __restrict__ B) {
    int lx = blockIdx.x;
    if (lx >= ncr) return;
    BrkCrest c = cr[lx];
    __shared__ int snb;
    if (threadIdx.x == 0) snb = 0;
    __syncthreads();
    if (!c.found) return;
    const int nx = g.nx, ny = g.ny;
    double dxg = g.x[1] - g.x[0], dyg = g.y[1] - g.y[0];
    int rx = (int)(g.R / dxg) + 2, ry = (int)(g.R / dyg) + 2;
    double xp = g.x[c.xNow], yp = g.y[c.yNow];
    int i0 = max(0, c.yNow - ry), i1 = min(ny - 1, c.yNow + ry);
    int j0 = max(0, c.xNow - rx), j1 = min(nx - 1, c.xNow + rx);
    int w = j1 - j0 + 1, cnt = w * (i1 - i0 + 1);
    int mine = 0;
    for (int p = threadIdx.x; p < cnt; p += blockDim.x) {
        int i = i0 + p / w, j = j0 + p % w;
        size_t q = (size_t)i * nx + j;
        if (owner[q] != lx) continue;
        if (!crest_char(g, i, j, xp, yp, c.theta, c.kwadran)) continue;
        double uu = u[q], vv = v[q], U = sqrt(uu * uu + vv * vv);
        double b = U >= 2 * c.Ustar ? 1.0 : (U <= c.Ustar ? 0.0 : U / c.Ustar - 1.0);
        B[q] = (T)b;
        mine++;
    }
    atomicAdd(&snb, mine);
    __syncthreads();
    if (threadIdx.x == 0) cr[lx].nb = snb;
}

template <typename T>
__global__ void k_fill(T* __restrict__ p, T val, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t s = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += s) p[i] = val;
}

// ---- eddy-viscosity source (fun_sbreaking) --------------------------------
// delphi = F^-1(P(dH1 + adjust(dH2)_eta)) = F^-1(P(dH1 + i F(g))) + chi A_eta,
// g = chi (-1)^x s_eta(y) the imaginary part of the adjust's inverse.
template <typename T, int NX, int NY>
__global__ void HW_XK_BOUNDS k_brk_g_x(Rhs2DArgs<T> a, const T* __restrict__ s_eta,
                                                              typename ctype<T>::type* X) {
    HW_XK_PROLOGUE
    C f[G::PP], t[G::PP];
    ph_load<T, NX, NY>(t, a.chi, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) {
        int pos = G::ppos(tid, q);
        int r = pos / G::L;
        T s = s_eta[y0 + (r < G::RR ? r : 0)];
        f[q] = cmk<C>(t[q].x * s, -t[q].y * s);
    }
    x_fwd_store<T, NX, NY>(f, X, y0, lds, twxs, tid);
}

template <typename T, int NX, int NY>
__global__ void __launch_bounds__((YGeo<T, NY>::NT)) k_brk_del_y(Rhs2DArgs<T> a, int flat, const typename ctype<T>::type* Xg,
                                                                typename ctype<T>::type* Xo) {
    HW_YK_PROLOGUE
    C Fg[G::PE], d[G::PE];
    y_load<T, NX, NY>(Fg, Xg, nullptr, col0, tid);
    y_fwd_regs<T, NX, NY>(Fg, lds, twys, tid);
    if (!flat) {
        y_load<T, NX, NY>(d, a.dH1, nullptr, col0, tid);
    } else {
        C uh[G::PE], vh[G::PE];
        y_load<T, NX, NY>(uh, a.u_hat, nullptr, col0, tid);
        y_load<T, NX, NY>(vh, a.v_hat, nullptr, col0, tid);
#pragma unroll
        for (int e = 0; e < G::PE; ++e) {
            int kx, ky; size_t gi;
            YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
            if (kx < NX / 2 + 1) {
                C s = cadd(cscale(uh[e], a.kxh[kx]), cscale(vh[e], a.ky[ky]));
                d[e] = cmul_i<-1>(cscale(s, a.c2d2g[gi]));
            } else {
                d[e] = cmk<C>((T)0, (T)0);
            }
        }
    }
#pragma unroll
    for (int e = 0; e < G::PE; ++e) d[e] = cadd(d[e], cmul_i<1>(Fg[e]));
    y_inv_regs<T, NX, NY>(d, lds, twys, col0, tid);
    y_store<T, NX, NY>(Xo, d, col0, tid);
}

// Flux = -B delb^2 (eta + D) delphi^2 (2d/fun_odesolver.c:995-1000)
template <typename T, int NX, int NY>
__global__ void HW_XK_BOUNDS k_brk_flux_x(Rhs2DArgs<T> a, const typename ctype<T>::type* aeta, const typename ctype<T>::type* Xd,
                                                                 const T* __restrict__ B, const T* __restrict__ depth, T delb2,
                                                                 T* __restrict__ flux) {
    HW_XK_PROLOGUE
    C dp[G::PP], A[G::PP], t[G::PP];
    x_inverse<T, NX, NY>(Xd, y0, lds, twxs, dp, sc, nullptr, tid);
    x_inverse<T, NX, NY>(aeta, y0, lds, twxs, A, sc, nullptr, tid);
    ph_load<T, NX, NY>(t, a.chi, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) dp[q] = pfma(t[q], A[q], dp[q]);
    ph_load<T, NX, NY>(t, a.eta_p, y0, tid);
    ph_load<T, NX, NY>(A, depth, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) {
        t[q] = cadd(t[q], A[q]);   // eta + D
    }
    ph_load<T, NX, NY>(A, B, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q)
        dp[q] = cmk<C>(-A[q].x * delb2 * t[q].x * dp[q].x * dp[q].x, -A[q].y * delb2 * t[q].y * dp[q].y * dp[q].y);
    ph_store<T, NX, NY>(flux, dp, y0, tid);
}

// Sb = clipped one-sided gradient of the flux / (eta + D)  (fun_gradient,
// 2d/functions.c:392-461, periodic differences)
template <typename T>
__global__ void k_brk_sb(const T* __restrict__ F, const T* __restrict__ eta, const T* __restrict__ depth, double dx, double dy,
                         int nx, int ny, T* __restrict__ sbx, T* __restrict__ sby) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y;
    if (j >= nx) return;
    size_t q = (size_t)i * nx + j;
    T f = F[q];
    T fx = (F[(size_t)i * nx + (j + 1) % nx] - f) / (T)dx;
    T fy = (F[(size_t)((i + 1) % ny) * nx + j] - f) / (T)dy;
    T bx = (f - F[(size_t)i * nx + (j + nx - 1) % nx]) / (T)dx;
    T by = (f - F[(size_t)((i + ny - 1) % ny) * nx + j]) / (T)dy;
    if (fx > 0) fx = 0;
    if (fy > 0) fy = 0;
    if (bx < 0) bx = 0;
    if (by < 0) by = 0;
    T H = eta[q] + depth[q];
    sbx[q] = (fx + bx) / H;
    sby[q] = (fy + by) / H;
}

// forward transforms of Sb_x, Sb_y into persistent half spectra
template <typename T, int NX, int NY>
__global__ void HW_XK_BOUNDS k_brk_sb_x(Rhs2DArgs<T> a, const T* sbx, const T* sby,
                                                               typename ctype<T>::type* X1, typename ctype<T>::type* X2) {
    HW_XK_PROLOGUE
    C f[G::PP];
    ph_load<T, NX, NY>(f, sbx, y0, tid);
    x_fwd_store<T, NX, NY>(f, X1, y0, lds, twxs, tid);
    ph_load<T, NX, NY>(f, sby, y0, tid);
    x_fwd_store<T, NX, NY>(f, X2, y0, lds, twxs, tid);
}

template <typename T, int NX, int NY>
__global__ void __launch_bounds__((YGeo<T, NY>::NT)) k_brk_sb_y(Rhs2DArgs<T> a, const typename ctype<T>::type* X1,
                                                               const typename ctype<T>::type* X2, typename ctype<T>::type* S1,
                                                               typename ctype<T>::type* S2) {
    HW_YK_PROLOGUE
    C v[G::PE];
    y_load<T, NX, NY>(v, X1, nullptr, col0, tid);
    y_fwd_regs<T, NX, NY>(v, lds, twys, tid);
    y_store<T, NX, NY>(S1, v, col0, tid);
    y_load<T, NX, NY>(v, X2, nullptr, col0, tid);
    y_fwd_regs<T, NX, NY>(v, lds, twys, tid);
    y_store<T, NX, NY>(S2, v, col0, tid);
}

// k_u += aal Sb_x^, k_v += aal Sb_y^ (current or stale, 2d/fun_odesolver.c:4599-4611)
template <typename T>
__global__ void k_add_sb(typename ctype<T>::type* __restrict__ ku, typename ctype<T>::type* __restrict__ kv,
                         const typename ctype<T>::type* __restrict__ sx, const typename ctype<T>::type* __restrict__ sy,
                         const T* __restrict__ aal, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t s = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += s) {
        T w = aal[i];
        ku[i] = cadd(ku[i], cscale(sx[i], w));
        kv[i] = cadd(kv[i], cscale(sy[i], w));
    }
}

}  // namespace hw
