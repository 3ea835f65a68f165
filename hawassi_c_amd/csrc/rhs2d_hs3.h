// rhs2d_hs3.h -- the third-order ("HS3") terms of the 2-D AB3 right-hand side.
//
// rhs_AB3_2DB_uv / rhs_AB3_2DF_uv (2d/fun_odesolver.c:4735-4843 / 4628-4733)
// add to the AB2 right-hand side, through fun_termshat_AB3_2DB_uv (:3899-4349)
// and fun_termshat_AB3_2DF_uv (:3512-3897),
//
//   eta eq.:  adjust(deluH3 + delvH3),
//             deluH3 + delvH3 = L(eta LLphi + eta^2 dgp w / 2)
//                               - (kx^2 F(eta^2 M0u w) + ky^2 F(eta^2 M0v w)) / 2
//   u, v eq.: - adjust(i kx F(dE3)), - adjust(i ky F(dE3)),
//             dE3 = Lphi (eta dgp + LLphi)
//
// with Lphi = M0u + M0v, dgp = F^-1(i (kx u^ + ky v^)),
// LLphi = F^-1(LeM0u^ + LeM0v^) (two-depth: divided by g), and L the HSS
// operator of the AB2 term (flat: the multiplier L2d; two-depth: HSSgen_2depth
// with the Om2 weights, divided by g).  (eta LLphi is eta LM0u + eta LM0v and
// dgp is dxu + dyv: the reference's four L-terms are one L of a sum.)
//
// fun_adjust is linear, so the HS3 terms are added to the AB2 adjust inputs
// before the output transforms, which also hands the breaking flux the
// delphiH that includes them (the reference sums deluH3 into delphiH_hat
// before fun_sbreaking, :4787-4812).
//
// Two-depth order:  Y1 X1 Y2 X3<H3> Y4 X5a  YA XB YC  X5b Y6<H3> X7 Y8
//   X5a  the AB2 HSS tails (first half of X5)
//   YA   LLphi^ = (F(acc1) + F(acc2)) / (2g) - dH2p - i kx Q1^ - i ky Q2^
//        (dH2p holds -i k Q^ - (Om- F(g- P) + Om+ F(g+ P)) / (2g)); dgp^
//   XB   the products A = eta LLphi + eta^2 dgp w / 2, g-+ A, B = eta^2 M0u w,
//        Cc = eta^2 M0v w, D = Lphi (eta dgp + LLphi)
//   YC   dH2p += (Om- F(g- A) + Om+ F(g+ A)) / (2g) - (kx^2 B^ + ky^2 Cc^) / 2;
//        inverse y of Om-+ A^ (the HSS tails); i kx D^, i ky D^ added to the
//        velocity adjust inputs
//   X5b  the HS3 HSS tail; the two velocity output transforms of X5
//   Y6   adds F(acc3) / (2g)
// Flat order:  y1 x2<H3> y3  YA XB YC  x4 y5, where YC adds the eta term and
// the velocity terms to the three adjust inputs of x4.
#pragma once
#include "rhs2d_kernels.h"

namespace hw {

// W slots of the HS3 kernels
template <bool FLAT> struct H3Slots;
template <> struct H3Slots<false> {
    enum : int { LL = 4, DGP = 5, A = 3, GMA = 6, GPA = 7, B = 9, CC = 10, D = 11, TM = 12, TP = 13,
                 AX = S_AX, AY = S_AY, GCA = 15, TC = 9 };
};
template <> struct H3Slots<true> {
    enum : int { LL = 3, DGP = 4, A = 6, GMA = -1, GPA = -1, B = 8, CC = 9, D = 10, TM = -1, TP = -1,
                 AX = F_AX, AY = F_AY, GCA = -1, TC = -1 };
};

// X5a: F(g- b- + g+ b+) of the two AB2 HSS tails (the first half of k2b_x5,
// with the second tail in slot S3_ACC2)
template <typename T, int NX, int NY>
__global__ void HW_XK_BOUNDS k3b_x5a(Rhs2DArgs<T> a) {
    HW_XK_PROLOGUE
    C f[G::PP], acc[G::PP], gm[G::PP], gp[G::PP];
    XSpec<T, NX> b;
    x_spec_load<T, NX, NY>(a.W[S_B1M], y0, b, tid);
    ph_load<T, NX, NY>(gm, a.gam_m, y0, tid);
    ph_load<T, NX, NY>(gp, a.gam_p, y0, tid);
    x_inv_roll<T, NX, NY>(b, a.W[S_B1P], y0, lds, twxs, f, sc, nullptr, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) acc[q] = pmul(gm[q], f[q]);
    x_inv_roll<T, NX, NY>(b, a.W[S_B2M], y0, lds, twxs, f, sc, nullptr, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) acc[q] = pfma(gp[q], f[q], acc[q]);
    x_fwd_store<T, NX, NY>(acc, a.W[S_ACC1], y0, lds, twxs, tid);
    x_inv_roll<T, NX, NY>(b, a.W[S_B2P], y0, lds, twxs, f, sc, nullptr, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) acc[q] = pmul(gm[q], f[q]);
    x_inv_roll<T, NX, NY>(b, nullptr, y0, lds, twxs, f, sc, nullptr, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) acc[q] = pfma(gp[q], f[q], acc[q]);
    x_fwd_store<T, NX, NY>(acc, a.W[S3_ACC2], y0, lds, twxs, tid);
}

// YA: LLphi^ and dgp^, inverted in y
template <typename T, int NX, int NY, bool FLAT>
__global__ void __launch_bounds__((YGeo<T, NY>::NT)) __attribute__((amdgpu_waves_per_eu(YGeo<T, NY>::WPE))) k3_ya(Rhs2DArgs<T> a) {
    HW_YK_PROLOGUE
    using SL = H3Slots<FLAT>;
    C L[G::PE], X[G::PE];
    const T ig = (T)1 / a.g;
    if constexpr (FLAT) {
        // LeM0u^ + LeM0v^ = L2d (F(eta M0u w) + F(eta M0v w))   (2d/fun_odesolver.c:3293-3297)
        y_load<T, NX, NY>(L, a.W[F_P1], nullptr, col0, tid);
        y_load<T, NX, NY>(X, a.W[F_P2], nullptr, col0, tid);
        y_fwd_regs<T, NX, NY>(L, lds, twys, tid);
        y_fwd_regs<T, NX, NY>(X, lds, twys, tid);
#pragma unroll
        for (int e = 0; e < G::PE; ++e) {
            int kx, ky; size_t gi;
            YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
            L[e] = kx < NX / 2 + 1 ? cscale(cadd(L[e], X[e]), a.l2d[gi]) : cmk<C>((T)0, (T)0);
        }
    } else {
        y_load<T, NX, NY>(L, a.W[S_ACC1], nullptr, col0, tid);
        y_load<T, NX, NY>(X, a.W[S3_ACC2], nullptr, col0, tid);
        y_fwd_regs<T, NX, NY>(L, lds, twys, tid);
        y_fwd_regs<T, NX, NY>(X, lds, twys, tid);
#pragma unroll
        for (int e = 0; e < G::PE; ++e) L[e] = cscale(cadd(L[e], X[e]), (T)0.5 * ig);
        y_load<T, NX, NY>(X, a.dH2p, nullptr, col0, tid);
#pragma unroll
        for (int e = 0; e < G::PE; ++e) L[e] = csub(L[e], X[e]);
        #pragma unroll
        for (int c = 0; c < 2; ++c) {
            y_load<T, NX, NY>(X, a.W[c == 0 ? S_Q1 : S_Q2], nullptr, col0, tid);
            y_fwd_regs<T, NX, NY>(X, lds, twys, tid);
#pragma unroll
            for (int e = 0; e < G::PE; ++e) {
                int kx, ky; size_t gi;
                YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
                if (kx < NX / 2 + 1) {
                    T kk = c == 0 ? a.kxh[kx] : a.ky[ky];
                    L[e] = csub(L[e], cmul_i<1>(cscale(X[e], kk)));
                } else {
                    L[e] = cmk<C>((T)0, (T)0);
                }
            }
        }
    }
    y_inv_regs<T, NX, NY>(L, lds, twys, col0, tid);
    y_store<T, NX, NY>(a.W[SL::LL], L, col0, tid);
    // dgp^ = i (kx u^ + ky v^)
    y_load<T, NX, NY>(X, a.u_hat, nullptr, col0, tid);
    y_load<T, NX, NY>(L, a.v_hat, nullptr, col0, tid);
#pragma unroll
    for (int e = 0; e < G::PE; ++e) {
        int kx, ky; size_t gi;
        YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
        X[e] = kx < NX / 2 + 1 ? cmul_i<1>(cadd(cscale(X[e], a.kxh[kx]), cscale(L[e], a.ky[ky]))) : cmk<C>((T)0, (T)0);
    }
    y_inv_regs<T, NX, NY>(X, lds, twys, col0, tid);
    y_store<T, NX, NY>(a.W[SL::DGP], X, col0, tid);
}

// XB: the physical products, forward in x (ND3: also g_mid A)
template <typename T, int NX, int NY, bool FLAT, bool ND3 = false>
__global__ void HW_XK_BOUNDS k3_xb(Rhs2DArgs<T> a) {
    HW_XK_PROLOGUE
    using SL = H3Slots<FLAT>;
    C ll[G::PP], dg[G::PP], e[G::PP], x[G::PP], t[G::PP];
    XSpec<T, NX> b;
    x_spec_load<T, NX, NY>(a.W[SL::LL], y0, b, tid);
    x_inv_roll<T, NX, NY>(b, a.W[SL::DGP], y0, lds, twxs, ll, sc, nullptr, tid);
    x_inv_roll<T, NX, NY>(b, nullptr, y0, lds, twxs, dg, sc, nullptr, tid);
    ph_load<T, NX, NY>(e, a.eta_p, y0, tid);
    // D = Lphi (eta dgp + LLphi), Lphi = M0u + M0v
    ph_load<T, NX, NY>(x, a.m0u_p, y0, tid);
    ph_load<T, NX, NY>(t, a.m0v_p, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) x[q] = pmul(cadd(x[q], t[q]), pfma(e[q], dg[q], ll[q]));
    x_fwd_store<T, NX, NY>(x, a.W[SL::D], y0, lds, twxs, tid);
    // A = eta (LLphi + eta dgp w / 2)
    ph_load<T, NX, NY>(t, a.wallchar, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) {
        C h = pmul(pmul(e[q], dg[q]), t[q]);
        x[q] = pmul(e[q], cmk<C>(ll[q].x + (T)0.5 * h.x, ll[q].y + (T)0.5 * h.y));
    }
    if constexpr (!FLAT) {
        ph_load<T, NX, NY>(ll, a.gam_m, y0, tid);
#pragma unroll
        for (int q = 0; q < G::PP; ++q) dg[q] = pmul(ll[q], x[q]);
        x_fwd_store<T, NX, NY>(dg, a.W[SL::GMA], y0, lds, twxs, tid);
        ph_load<T, NX, NY>(ll, a.gam_p, y0, tid);
#pragma unroll
        for (int q = 0; q < G::PP; ++q) dg[q] = pmul(ll[q], x[q]);
        x_fwd_store<T, NX, NY>(dg, a.W[SL::GPA], y0, lds, twxs, tid);
        if constexpr (ND3) {
            ph_load<T, NX, NY>(ll, a.gam_c, y0, tid);
#pragma unroll
            for (int q = 0; q < G::PP; ++q) dg[q] = pmul(ll[q], x[q]);
            x_fwd_store<T, NX, NY>(dg, a.W[SL::GCA], y0, lds, twxs, tid);
        }
    }
    x_fwd_store<T, NX, NY>(x, a.W[SL::A], y0, lds, twxs, tid);
    // B = eta^2 M0u w, Cc = eta^2 M0v w
#pragma unroll
    for (int q = 0; q < G::PP; ++q) t[q] = pmul(pmul(e[q], e[q]), t[q]);
    ph_load<T, NX, NY>(x, a.m0u_p, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) x[q] = pmul(t[q], x[q]);
    x_fwd_store<T, NX, NY>(x, a.W[SL::B], y0, lds, twxs, tid);
    ph_load<T, NX, NY>(x, a.m0v_p, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) x[q] = pmul(t[q], x[q]);
    x_fwd_store<T, NX, NY>(x, a.W[SL::CC], y0, lds, twxs, tid);
}

// YC: the eta term (two-depth: into dH2p, with the HSS tails inverted in y;
// flat: inverted in y and added to the eta adjust input) and the velocity
// terms added to the velocity adjust inputs
template <typename T, int NX, int NY, bool FLAT, bool ND3 = false>
__global__ void __launch_bounds__((YGeo<T, NY>::NT)) __attribute__((amdgpu_waves_per_eu(YGeo<T, NY>::WPE))) k3_yc(Rhs2DArgs<T> a) {
    HW_YK_PROLOGUE
    using SL = H3Slots<FLAT>;
    C P[G::PE], Q[G::PE], R[G::PE];
    const T ig = (T)1 / a.g;
    if constexpr (FLAT) {
        y_load<T, NX, NY>(P, a.W[SL::A], nullptr, col0, tid);
        y_fwd_regs<T, NX, NY>(P, lds, twys, tid);
#pragma unroll
        for (int e = 0; e < G::PE; ++e) {
            int kx, ky; size_t gi;
            YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
            P[e] = kx < NX / 2 + 1 ? cscale(P[e], a.l2d[gi]) : cmk<C>((T)0, (T)0);
        }
    } else {
        y_load<T, NX, NY>(P, a.W[SL::GMA], nullptr, col0, tid);
        y_load<T, NX, NY>(Q, a.W[SL::GPA], nullptr, col0, tid);
        y_fwd_regs<T, NX, NY>(P, lds, twys, tid);
        y_fwd_regs<T, NX, NY>(Q, lds, twys, tid);
#pragma unroll
        for (int e = 0; e < G::PE; ++e) {
            int kx, ky; size_t gi;
            YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
            P[e] = kx < NX / 2 + 1 ? cscale(cadd(cscale(P[e], a.om2_m[gi]), cscale(Q[e], a.om2_p[gi])), (T)0.5 * ig)
                                   : cmk<C>((T)0, (T)0);
        }
        if constexpr (ND3) {
            y_load<T, NX, NY>(Q, a.W[SL::GCA], nullptr, col0, tid);
            y_fwd_regs<T, NX, NY>(Q, lds, twys, tid);
#pragma unroll
            for (int e = 0; e < G::PE; ++e) {
                int kx, ky; size_t gi;
                YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
                if (kx < NX / 2 + 1) P[e] = cadd(P[e], cscale(Q[e], (T)0.5 * ig * a.om2_c[gi]));
            }
        }
    }
    // - (kx^2 B^ + ky^2 Cc^) / 2
    #pragma unroll
    for (int c = 0; c < 2; ++c) {
        y_load<T, NX, NY>(Q, a.W[c == 0 ? SL::B : SL::CC], nullptr, col0, tid);
        y_fwd_regs<T, NX, NY>(Q, lds, twys, tid);
#pragma unroll
        for (int e = 0; e < G::PE; ++e) {
            int kx, ky; size_t gi;
            YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
            if (kx < NX / 2 + 1) {
                T kk = c == 0 ? a.kxh[kx] : a.ky[ky];
                P[e] = csub(P[e], cscale(Q[e], (T)0.5 * kk * kk));
            }
        }
    }
    if constexpr (FLAT) {
        y_inv_regs<T, NX, NY>(P, lds, twys, col0, tid);
        y_load<T, NX, NY>(Q, a.W[F_AETA], nullptr, col0, tid);
#pragma unroll
        for (int e = 0; e < G::PE; ++e) P[e] = cadd(P[e], Q[e]);
        y_store<T, NX, NY>(a.W[F_AETA], P, col0, tid);
    } else {
        y_load<T, NX, NY>(Q, a.dH2p, nullptr, col0, tid);
#pragma unroll
        for (int e = 0; e < G::PE; ++e) P[e] = cadd(P[e], Q[e]);
        y_store<T, NX, NY>(a.dH2p, P, col0, tid);
        // tails: inverse y of Om2-+ A^
        y_load<T, NX, NY>(P, a.W[SL::A], nullptr, col0, tid);
        y_fwd_regs<T, NX, NY>(P, lds, twys, tid);
        #pragma unroll
        for (int s = 0; s < (ND3 ? 3 : 2); ++s) {
#pragma unroll
            for (int e = 0; e < G::PE; ++e) {
                int kx, ky; size_t gi;
                YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
                const T* om = s == 0 ? a.om2_m : s == 1 ? a.om2_p : a.om2_c;
                Q[e] = kx < NX / 2 + 1 ? cscale(P[e], om[gi]) : cmk<C>((T)0, (T)0);
            }
            y_inv_regs<T, NX, NY>(Q, lds, twys, col0, tid);
            y_store<T, NX, NY>(a.W[s == 0 ? SL::TM : s == 1 ? SL::TP : SL::TC], Q, col0, tid);
        }
    }
    // velocity adjust inputs += i kx D^, i ky D^
    y_load<T, NX, NY>(P, a.W[SL::D], nullptr, col0, tid);
    y_fwd_regs<T, NX, NY>(P, lds, twys, tid);
    #pragma unroll
    for (int s = 0; s < 2; ++s) {
#pragma unroll
        for (int e = 0; e < G::PE; ++e) {
            int kx, ky; size_t gi;
            YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
            T kk = (kx < NX / 2 + 1) ? (s == 0 ? a.kxh[kx] : a.ky[ky]) : (T)0;
            Q[e] = cmul_i<1>(cscale(P[e], kk));
        }
        y_inv_regs<T, NX, NY>(Q, lds, twys, col0, tid);
        y_load<T, NX, NY>(R, a.W[s == 0 ? SL::AX : SL::AY], nullptr, col0, tid);
#pragma unroll
        for (int e = 0; e < G::PE; ++e) Q[e] = cadd(Q[e], R[e]);
        y_store<T, NX, NY>(a.W[s == 0 ? SL::AX : SL::AY], Q, col0, tid);
    }
}

// X5b: F(g- b3- + g+ b3+) of the HS3 tail (ND3: + g_mid b3_mid); the two
// velocity output transforms
template <typename T, int NX, int NY, bool ND3 = false>
__global__ void HW_XK_BOUNDS k3b_x5b(Rhs2DArgs<T> a) {
    HW_XK_PROLOGUE
    using SL = H3Slots<false>;
    C f[G::PP], acc[G::PP], gg[G::PP];
    XSpec<T, NX> b;
    x_spec_load<T, NX, NY>(a.W[SL::TM], y0, b, tid);
    x_inv_roll<T, NX, NY>(b, a.W[SL::TP], y0, lds, twxs, f, sc, nullptr, tid);
    ph_load<T, NX, NY>(gg, a.gam_m, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) acc[q] = pmul(gg[q], f[q]);
    x_inv_roll<T, NX, NY>(b, ND3 ? a.W[SL::TC] : a.W[S_AX], y0, lds, twxs, f, sc, nullptr, tid);
    ph_load<T, NX, NY>(gg, a.gam_p, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) acc[q] = pfma(gg[q], f[q], acc[q]);
    if constexpr (ND3) {
        x_inv_roll<T, NX, NY>(b, a.W[S_AX], y0, lds, twxs, f, sc, nullptr, tid);
        ph_load<T, NX, NY>(gg, a.gam_c, y0, tid);
#pragma unroll
        for (int q = 0; q < G::PP; ++q) acc[q] = pfma(gg[q], f[q], acc[q]);
    }
    x_fwd_store<T, NX, NY>(acc, a.W[S3_ACC3], y0, lds, twxs, tid);
    x_vel_out<T, NX, NY>(a, b, a.W[S_AY], a.u_p, a.sbx, S_ZU, y0, lds, twxs, nyq, tid);
    x_vel_out<T, NX, NY>(a, b, nullptr, a.v_p, a.sby, S_ZV, y0, lds, twxs, nyq, tid);
}

}  // namespace hw
