// rhs2d_3depth.h -- the third reference depth of HSSgen_3depth.
//
// With ninterp = 3 the reference evaluates every HSS operator over three
// reference depths (HSSgen_3depth, 2d/fun_odesolver.c:588-697):
//   HSS(phi) = sum_j (F(g_j F^-1(O_j phi^)) + O_j F(g_j phi)) / 2,
// j = min, mid, plus.  The operator is a sum over depths, so the middle
// depth's share is added to the two-depth pipeline by these kernels, each
// at the point where the pipeline still holds the quantity it corrects:
//
//   after Y2  (M0 = -i k HSS_C2(u) / g, 2d/fun_odesolver.c:2914-2920)
//     D1 (y)  inverse y of C2_mid u^, C2_mid v^
//     D2 (x)  F(g_mid F^-1(C2_mid u^)), F(g_mid u) (and v)
//     D3 (y)  dM0u^ = -i kx (F(tail) + C2_mid F(g_mid u)) / (2g): added to
//             dH1 and, inverted in y, to the M0u / M0v inputs of X3
//   after Y4  (LeM0 = HSS_Om2(eta M0 w), :3459-3465; one operator of the sum
//             P = eta (M0u + M0v) w, by linearity)
//     E1 (x)  F(g_mid P)
//     E2 (y)  dH2p -= Om2_mid F(g_mid P) / (2g); inverse y of Om2_mid P^
//   after X5 / X5a
//     E3 (x)  F(g_mid F^-1(Om2_mid P^)), added to the first HSS tail sum
//             (k_add_x), which Y6 (and YA for HS3) transform in y
// The HS3 kernels take the middle depth of their own HSS through their ND3
// variants.
#pragma once
#include "rhs2d_hs3.h"

namespace hw {

enum : int { D3_IU = 2, D3_IV = 3, D3_TU = 4, D3_TV = 5, D3_GU = 6, D3_GV = 7, E3_GC = 9, E3_TC = 15, E3_ACC = 10 };

template <typename T, int NX, int NY>
__global__ void __launch_bounds__((YGeo<T, NY>::NT)) __attribute__((amdgpu_waves_per_eu(YGeo<T, NY>::WPE))) k3d_y1(Rhs2DArgs<T> a) {
    HW_YK_PROLOGUE
    C v[G::PE];
    #pragma unroll
    for (int c = 0; c < 2; ++c) {
        y_load<T, NX, NY>(v, c == 0 ? a.u_hat : a.v_hat, a.csq_c, col0, tid);
        y_inv_regs<T, NX, NY>(v, lds, twys, col0, tid);
        y_store<T, NX, NY>(a.W[c == 0 ? D3_IU : D3_IV], v, col0, tid);
    }
}

template <typename T, int NX, int NY>
__global__ void HW_XK_BOUNDS k3d_x2(Rhs2DArgs<T> a) {
    HW_XK_PROLOGUE
    C f[G::PP], g[G::PP];
    XSpec<T, NX> b;
    ph_load<T, NX, NY>(g, a.gam_c, y0, tid);
    x_spec_load<T, NX, NY>(a.W[D3_IU], y0, b, tid);
    x_inv_roll<T, NX, NY>(b, a.W[D3_IV], y0, lds, twxs, f, sc, nullptr, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) f[q] = pmul(g[q], f[q]);
    x_fwd_store<T, NX, NY>(f, a.W[D3_TU], y0, lds, twxs, tid);
    x_inv_roll<T, NX, NY>(b, nullptr, y0, lds, twxs, f, sc, nullptr, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) f[q] = pmul(g[q], f[q]);
    x_fwd_store<T, NX, NY>(f, a.W[D3_TV], y0, lds, twxs, tid);
    #pragma unroll
    for (int c = 0; c < 2; ++c) {
        ph_load<T, NX, NY>(f, c == 0 ? a.u_p : a.v_p, y0, tid);
#pragma unroll
        for (int q = 0; q < G::PP; ++q) f[q] = pmul(g[q], f[q]);
        x_fwd_store<T, NX, NY>(f, a.W[c == 0 ? D3_GU : D3_GV], y0, lds, twxs, tid);
    }
}

template <typename T, int NX, int NY>
__global__ void __launch_bounds__((YGeo<T, NY>::NT)) __attribute__((amdgpu_waves_per_eu(YGeo<T, NY>::WPE))) k3d_y3(Rhs2DArgs<T> a) {
    HW_YK_PROLOGUE
    C A[G::PE], B[G::PE], dsum[G::PE];
    const T ig = (T)1 / a.g;
    #pragma unroll
    for (int c = 0; c < 2; ++c) {
        y_load<T, NX, NY>(A, a.W[c == 0 ? D3_TU : D3_TV], nullptr, col0, tid);
        y_load<T, NX, NY>(B, a.W[c == 0 ? D3_GU : D3_GV], nullptr, col0, tid);
        y_fwd_regs<T, NX, NY>(A, lds, twys, tid);
        y_fwd_regs<T, NX, NY>(B, lds, twys, tid);
#pragma unroll
        for (int e = 0; e < G::PE; ++e) {
            int kx, ky; size_t gi;
            YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
            C d = cmk<C>((T)0, (T)0);
            if (kx < NX / 2 + 1) {
                C hs = cscale(cadd(A[e], cscale(B[e], a.csq_c[gi])), (T)0.5);
                d = cmul_i<-1>(cscale(hs, (c == 0 ? a.kxh[kx] : a.ky[ky]) * ig));
            }
            A[e] = d;
            dsum[e] = c == 0 ? d : cadd(dsum[e], d);
        }
        y_inv_regs<T, NX, NY>(A, lds, twys, col0, tid);
        y_load<T, NX, NY>(B, a.W[c == 0 ? S_M0U : S_M0V], nullptr, col0, tid);
#pragma unroll
        for (int e = 0; e < G::PE; ++e) A[e] = cadd(A[e], B[e]);
        y_store<T, NX, NY>(a.W[c == 0 ? S_M0U : S_M0V], A, col0, tid);
    }
    y_load<T, NX, NY>(B, a.dH1, nullptr, col0, tid);
#pragma unroll
    for (int e = 0; e < G::PE; ++e) dsum[e] = cadd(dsum[e], B[e]);
    y_store<T, NX, NY>(a.dH1, dsum, col0, tid);
}

template <typename T, int NX, int NY>
__global__ void HW_XK_BOUNDS k3d_x4(Rhs2DArgs<T> a) {
    HW_XK_PROLOGUE
    C m[G::PP], t[G::PP];
    ph_load<T, NX, NY>(m, a.m0u_p, y0, tid);
    ph_load<T, NX, NY>(t, a.m0v_p, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) m[q] = cadd(m[q], t[q]);
    ph_load<T, NX, NY>(t, a.eta_p, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) m[q] = pmul(m[q], t[q]);
    ph_load<T, NX, NY>(t, a.wallchar, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) m[q] = pmul(m[q], t[q]);
    ph_load<T, NX, NY>(t, a.gam_c, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) m[q] = pmul(m[q], t[q]);
    x_fwd_store<T, NX, NY>(m, a.W[E3_GC], y0, lds, twxs, tid);
}

template <typename T, int NX, int NY>
__global__ void __launch_bounds__((YGeo<T, NY>::NT)) __attribute__((amdgpu_waves_per_eu(YGeo<T, NY>::WPE))) k3d_y5(Rhs2DArgs<T> a) {
    HW_YK_PROLOGUE
    C P[G::PE], Q[G::PE], R[G::PE];
    const T ig = (T)1 / a.g;
    y_load<T, NX, NY>(P, a.W[S_P1], nullptr, col0, tid);
    y_load<T, NX, NY>(Q, a.W[S_P2], nullptr, col0, tid);
#pragma unroll
    for (int e = 0; e < G::PE; ++e) P[e] = cadd(P[e], Q[e]);
    y_fwd_regs<T, NX, NY>(P, lds, twys, tid);
    y_load<T, NX, NY>(Q, a.W[E3_GC], nullptr, col0, tid);
    y_fwd_regs<T, NX, NY>(Q, lds, twys, tid);
    y_load<T, NX, NY>(R, a.dH2p, nullptr, col0, tid);
#pragma unroll
    for (int e = 0; e < G::PE; ++e) {
        int kx, ky; size_t gi;
        YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
        if (kx < NX / 2 + 1) {
            R[e] = csub(R[e], cscale(Q[e], (T)0.5 * ig * a.om2_c[gi]));
            Q[e] = cscale(P[e], a.om2_c[gi]);
        } else {
            Q[e] = cmk<C>((T)0, (T)0);
        }
    }
    y_store<T, NX, NY>(a.dH2p, R, col0, tid);
    y_inv_regs<T, NX, NY>(Q, lds, twys, col0, tid);
    y_store<T, NX, NY>(a.W[E3_TC], Q, col0, tid);
}

template <typename T, int NX, int NY>
__global__ void HW_XK_BOUNDS k3d_x6(Rhs2DArgs<T> a) {
    HW_XK_PROLOGUE
    C f[G::PP], g[G::PP];
    XSpec<T, NX> b;
    x_spec_load<T, NX, NY>(a.W[E3_TC], y0, b, tid);
    x_inv_roll<T, NX, NY>(b, nullptr, y0, lds, twxs, f, sc, nullptr, tid);
    ph_load<T, NX, NY>(g, a.gam_c, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) f[q] = pmul(g[q], f[q]);
    x_fwd_store<T, NX, NY>(f, a.W[E3_ACC], y0, lds, twxs, tid);
}

// dst += src over n complex elements
template <typename C>
__global__ void k_add_x(C* __restrict__ dst, const C* __restrict__ src, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t s = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += s) dst[i] = cadd(dst[i], src[i]);
}

}  // namespace hw
