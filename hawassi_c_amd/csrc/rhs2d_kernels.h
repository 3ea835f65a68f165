// rhs2d_kernels.h -- fused HIP kernels of the 2-D AB2 ("HS2") right-hand side.
//
// The reference evaluates rhs_AB2_2DB_uv (2d/fun_odesolver.c:4542-4624) with
// 51-55 full complex FFTs.  Here the same operator is evaluated on the half
// spectrum with every pointwise product, multiplier, mask, projection and the
// linear combinations folded into the prologue/epilogue of 9 FFT passes
// (x = row passes, y = column passes):
//
//   Y1  inverse y of eta^, u^, v^, C2-(k) u^, C2+(k) u^, C2-(k) v^, C2+(k) v^
//   X1  inverse x of those; forward x of g-u, g+u, g-a-u + g+a+u (and v)     [HSSgen_2depth]
//   Y2  forward y; dH1 = -i/g (kx HSS(u) + ky HSS(v)); inverse y of M0u, M0v
//   X3  inverse x of M0u, M0v; forward x of eta M0u w, eta u w, eta M0v w,
//       eta v w, E w and g+-(eta M0 w)                                      [AB2 products]
//   Y4  forward y; partial dH2; inverse y of Om2+-(k) F(eta M0 w), i kx E^, i ky E^
//   X5  inverse x; forward x of g-b- + g+b+; adjust of i kx E^, i ky E^ folded
//       with damping into the u and v output transforms                    [fun_adjust]
//   Y6  forward y; dH2 = dH2part - (F(acc_u)+F(acc_v))/(2g); inverse y of dH2;
//       du^/dt, dv^/dt written
//   X7  inverse x of dH2 (adjust), source(t) and damping folded into the
//       eta output transform                                                [fun_source]
//   Y8  forward y; deta^/dt written
//
// Flat bottom (rhs_AB2_2DF_uv, 2d/fun_odesolver.c:4458-4540) uses 5 passes.
#pragma once
#include "hw_fft.h"

namespace hw {

template <typename T> struct Rhs2DArgs {
    using C = typename ctype<T>::type;
    // RHS input, half spectra [kx][ky]
    const C *eta_hat, *u_hat, *v_hat;
    // spectral multipliers, half layout [kx][ky]
    const T *csq_m, *csq_p, *om2_m, *om2_p, *aal, *c2d2g, *l2d;
    const T *kxh, *ky;   // kx for kx=0..NX/2, ky for all ky
    // physical coefficients [y][x]
    const T *gam_m, *gam_p, *wallchar, *fbdy7, *chi, *new_gam, *skew_gam;
    // source line values at t (per row for vertical, per column for horizontal)
    const T *src_sig, *src_skew;
    int src_on, vertical, propagation;
    // breaking source, physical [y][x] (null when breaking is off)
    const T *sbx, *sby;
    // physical work fields
    T *eta_p, *u_p, *v_p;
    // x-intermediates [kx][y] / spectral work
    C *W[16];
    C *dH1;     // dH1 = duH1 + dvH1 (2DB), kept for the eta output
    C *dH2p;    // partial dH2
    // outputs: d/dt of eta^, u^, v^, half spectra
    C *k_eta, *k_u, *k_v;
    const C *twx, *twy;
    T g;
    T* s_eta_out;   // per-row imaginary Nyquist part of the eta adjust (read by breaking)
    C* eta_col0;    // when set: the pre-mask eta derivative of column kx = 0, [ky]
    T *m0u_p, *m0v_p;   // physical M0u, M0v (HS3 and three reference depths)
    // third reference depth (HSSgen_3depth): gamma_mid [y][x], C^2_mid and
    // Om^2_mid [kx][ky]; null for two depths
    const T *gam_c, *csq_c, *om2_c;
    // The host step loop hands the first y-pass two side jobs that would
    // otherwise be launches of their own:
    // * the influx line signals at t (natural cubic splines, as k_src_eval):
    //   when sp_n > 0, into src_sig_w / src_skew_w;
    // * zero_cnt: when set, the two peak-search counters are zeroed.
    const double *sp_y, *sp_c, *sk_y, *sk_c;
    int sp_n, sp_idx;
    double sp_dx, sp_hh;
    T *src_sig_w, *src_skew_w;
    int* zero_cnt;
    // Several states of one model in one launch (grid.z = member): member z's
    // input / output spectra, work slots, physical fields and per-row values
    // sit these many elements after member 0's.
    size_t zs_state, zs_work, zs_phys, zs_row;
    __device__ __forceinline__ void shift(int z) {
        if (z == 0) return;
        const size_t s = (size_t)z * zs_state, w = (size_t)z * zs_work, p = (size_t)z * zs_phys;
        eta_hat += s; u_hat += s; v_hat += s;
        k_eta += s; k_u += s; k_v += s;
#pragma unroll
        for (int i = 0; i < 16; ++i) W[i] += w;
        dH1 += w; dH2p += w;
        eta_p += p; u_p += p; v_p += p;
        if (m0u_p) { m0u_p += p; m0v_p += p; }
        if (s_eta_out) s_eta_out += (size_t)z * zs_row;
    }
};

template <typename T, int NX, int NY> struct Geo2 {
    using C = typename ctype<T>::type;
    using XG = XGeo<T, NX>;
    using YG = YGeo<T, NY>;
    static constexpr int NXH = NX / 2 + 1;
    static constexpr int L = NX / 2;
};

// ---------------------------------------------------------------- helpers --
template <typename T, int NX, int NY>
__device__ __forceinline__ void ph_load(carr<T, XGeo<T, NX>::PP>& v, const T* __restrict__ a, int y0, int tid) {
    using G = XGeo<T, NX>;
    using C = typename ctype<T>::type;
    const C* a2 = reinterpret_cast<const C*>(a) + (size_t)y0 * G::L;
#pragma unroll
    for (int q = 0; q < G::PP; ++q) {
        int pos = G::ppos(tid, q);
        if (G::NPAIR % G::NT == 0 || pos < G::NPAIR) v[q] = a2[pos];
    }
}

template <typename T, int NX, int NY>
__device__ __forceinline__ void ph_store(T* __restrict__ a, const carr<T, XGeo<T, NX>::PP>& v, int y0, int tid) {
    using G = XGeo<T, NX>;
    using C = typename ctype<T>::type;
    C* a2 = reinterpret_cast<C*>(a) + (size_t)y0 * G::L;
#pragma unroll
    for (int q = 0; q < G::PP; ++q) {
        int pos = G::ppos(tid, q);
        if (G::NPAIR % G::NT == 0 || pos < G::NPAIR) a2[pos] = v[q];
    }
}

// elementwise helpers on pair registers (each C holds two real samples)
template <typename C> __device__ __forceinline__ C pmul(C a, C b) { return cmk<C>(a.x * b.x, a.y * b.y); }
template <typename C> __device__ __forceinline__ C pfma(C a, C b, C c) { return cmk<C>(a.x * b.x + c.x, a.y * b.y + c.y); }

// forward x of a real field in registers, stored transposed into X
template <typename T, int NX, int NY>
__device__ __forceinline__ void x_fwd_store(const carr<T, XGeo<T, NX>::PP>& f, typename ctype<T>::type* X,
                                            int y0, typename ctype<T>::type* lds, const typename ctype<T>::type* twx, int tid) {
    using G = XGeo<T, NX>;
    const auto* O = x_fwd_to_lds<T, NX>(f, lds, twx, tid);
#pragma unroll
    for (int q = 0; q < G::PS; ++q)
        if (x_slot_ok<T, NX>(q, tid)) x_store_slot<T, NX, NY>(X, y0, q, x_post<T, NX>(O, twx, q, tid), tid);
}

// forward x of f1 + i f2 (both real): X = F(f1) + i F(f2)
template <typename T, int NX, int NY>
__device__ __forceinline__ void x_fwd2_store(const carr<T, XGeo<T, NX>::PP>& f1,
                                             const carr<T, XGeo<T, NX>::PP>& f2, typename ctype<T>::type* X,
                                             int y0, typename ctype<T>::type* lds, const typename ctype<T>::type* twx, int tid) {
    using G = XGeo<T, NX>;
    XSpec<T, NX> s1, s2;
    x_forward_regs<T, NX>(f1, s1, lds, twx, tid);
    const auto* O = x_fwd_to_lds<T, NX>(f2, lds, twx, tid);
#pragma unroll
    for (int q = 0; q < G::PS; ++q)
        if (x_slot_ok<T, NX>(q, tid))
            x_store_slot<T, NX, NY>(X, y0, q, cadd(s1[q], cmul_i<1>(x_post<T, NX>(O, twx, q, tid))), tid);
}

// forward x of f1 + i f2 where f2 = c (-1)^x s(y) (c real per point, s per
// row): the adjust operator's imaginary part.  f2 is formed after F(f1) so
// that only c stays live during the first transform.
template <typename T, int NX, int NY>
__device__ __forceinline__ void x_fwd_adj_store(const carr<T, XGeo<T, NX>::PP>& f1, const carr<T, XGeo<T, NX>::PP>& c,
                                                T sgn, const T* nyq, typename ctype<T>::type* X, int y0,
                                                typename ctype<T>::type* lds, const typename ctype<T>::type* twx, int tid) {
    using G = XGeo<T, NX>;
    using C = typename ctype<T>::type;
    XSpec<T, NX> s1;
    x_forward_regs<T, NX>(f1, s1, lds, twx, tid);
    C f2[G::PP];
#pragma unroll
    for (int q = 0; q < G::PP; ++q) {
        const int r = G::ppos(tid, q) / G::L;
        const T sv = sgn * nyq[r < G::RR ? r : 0];
        f2[q] = cmk<C>(c[q].x * sv, -c[q].y * sv);
    }
    const auto* O = x_fwd_to_lds<T, NX>(f2, lds, twx, tid);
#pragma unroll
    for (int q = 0; q < G::PS; ++q)
        if (x_slot_ok<T, NX>(q, tid))
            x_store_slot<T, NX, NY>(X, y0, q, cadd(s1[q], cmul_i<1>(x_post<T, NX>(O, twx, q, tid))), tid);
}

// Inverse x of the field held in b; b is refilled with `next` (when set)
// while the transform runs, so one spectral buffer keeps the loads of the
// following field in flight.
template <typename T, int NX, int NY>
__device__ __forceinline__ void x_inv_roll(XSpec<T, NX>& b, const typename ctype<T>::type* next, int y0,
                                           typename ctype<T>::type* lds, const typename ctype<T>::type* twl,
                                           carr<T, XGeo<T, NX>::PP>& out, T scale, T* nyq_im, int tid) {
    x_inv_stage<T, NX>(b, lds, tid);
    if (next) x_spec_load<T, NX, NY>(next, y0, b, tid);
    x_inv_finish<T, NX>(lds, twl, out, scale, nyq_im, tid);
}

// y-pass column element load with optional real multiplier
template <typename T, int NX, int NY>
__device__ __forceinline__ void y_load(carr<T, YGeo<T, NY>::PE>& v, const typename ctype<T>::type* __restrict__ src,
                                       const T* __restrict__ mul, int col0, int tid) {
    using G = YGeo<T, NY>;
    using C = typename ctype<T>::type;
    constexpr int NXH = NX / 2 + 1;
#pragma unroll
    for (int e = 0; e < G::PE; ++e) {
        int idx = G::eidx(tid, e);
        int c = idx / NY;
        bool ok = (G::NE % G::NT == 0 || idx < G::NE) && (col0 + c < NXH);
        if (ok) {
            size_t gi = (size_t)col0 * NY + idx;
            C x = src[gi];
            if (mul) x = cscale(x, mul[gi]);
            v[e] = x;
        } else {
            v[e] = cmk<C>((T)0, (T)0);
        }
    }
}

template <typename T, int NX, int NY>
__device__ __forceinline__ void y_store(typename ctype<T>::type* __restrict__ dst, const carr<T, YGeo<T, NY>::PE>& v,
                                        int col0, int tid) {
    using G = YGeo<T, NY>;
    constexpr int NXH = NX / 2 + 1;
#pragma unroll
    for (int e = 0; e < G::PE; ++e) {
        int idx = G::eidx(tid, e);
        int c = idx / NY;
        bool ok = (G::NE % G::NT == 0 || idx < G::NE) && (col0 + c < NXH);
        if (ok) dst[(size_t)col0 * NY + idx] = v[e];
    }
}

// the first y-pass's side jobs (see Rhs2DArgs): the peak-search
// counters and the influx line signals at t
template <typename T, int NT>
__device__ __forceinline__ void y1_side_jobs(const Rhs2DArgs<T>& a, int tid) {
    if (a.zero_cnt && blockIdx.x == 0 && tid < 2) a.zero_cnt[tid] = 0;
    if (a.sp_n > 0) {
        for (int j = blockIdx.x * NT + tid; j < a.sp_n; j += gridDim.x * NT) {
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const double* Y = s == 0 ? a.sp_y : a.sk_y;
                const double* Cc = s == 0 ? a.sp_c : a.sk_c;
                const double hh = a.sp_hh, dx = a.sp_dx;
                const size_t i0 = (size_t)a.sp_idx * a.sp_n + j, i1 = i0 + a.sp_n;
                const double y0 = Y[i0], y1 = Y[i1], ci = Cc[i0], cip = Cc[i1];
                const double b = (y1 - y0) / hh - hh * (cip + 2.0 * ci) / 3.0;
                const double d = (cip - ci) / (3.0 * hh);
                const double v = y0 + dx * (b + dx * (ci + dx * d));
                if (s == 0) a.src_sig_w[j] = (T)v; else a.src_skew_w[j] = (T)v;
            }
        }
    }
}

// per-element spectral coordinates of this thread's y-pass element e
template <typename T, int NX, int NY> struct YIdx {
    __device__ __forceinline__ static void get(int col0, int tid, int e, int& kx, int& ky, size_t& gi) {
        using G = YGeo<T, NY>;
        int idx = G::eidx(tid, e);
        kx = col0 + idx / NY;
        ky = idx % NY;
        gi = (size_t)col0 * NY + idx;
    }
};

// The eta derivative's last step: aal times the pre-mask value.  In a call
// that runs breaking, fun_sbreaking has projected delphiH in place before
// the reference forms the eta derivative (2d/fun_odesolver.c:990,
// fftwcomplex_2d_sym), which on the stored half means: column 0's rows
// ky > ny/2 are the conjugates of rows ny - ky, and Im = 0 at ky = 0, ny/2 of
// columns 0 and nx/2 (y_project_regs, through LDS for column 0).  The
// source and damping terms folded into the eta derivative are transforms of
// real fields, hence already Hermitian along column 0 and real at those
// bins, so projecting the whole pre-mask sum is projecting delphiH alone.
template <typename T, int NX, int NY>
__device__ __forceinline__ void y_eta_mask(carr<T, YGeo<T, NY>::PE>& Z, const Rhs2DArgs<T>& a,
                                           typename ctype<T>::type* lds, int col0, int tid) {
    using G = YGeo<T, NY>;
    if (a.eta_col0) {            // block-uniform: breaking runs in this call
        __syncthreads();         // the transform's last reads of lds are done
        y_project_regs<T, NX, NY>(Z, lds, col0, tid);
    }
#pragma unroll
    for (int e = 0; e < G::PE; ++e) {
        int kx, ky; size_t gi;
        YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
        if (kx < NX / 2 + 1) Z[e] = cscale(Z[e], a.aal[gi]);
    }
}

// ========================================================== 2DB kernels ==
#define HW_XK_PROLOGUE                                                   \
    a.shift((int)blockIdx.z);                                            \
    HW_XK_PROLOGUE_TW(a.twx)
#define HW_XK_PROLOGUE_TW(TWSRC)                                         \
    using C = typename ctype<T>::type;                                   \
    using G = XGeo<T, NX>;                                               \
    __shared__ C lds[G::LDS_ELEMS];                                      \
    __shared__ T nyq[G::RR];                                             \
    __shared__ C twxs_l[G::TWL ? G::TWN : 1];                            \
    const C* twxs = G::TWL ? twxs_l : (TWSRC);                           \
    const int tid = threadIdx.x;                                         \
    const int y0 = xcd_remap(blockIdx.x, gridDim.x) * G::RR;             \
    const T sc = (T)1 / ((T)NX * (T)NY);                                 \
    if (G::TWL)                                                          \
        for (int i_ = tid; i_ < G::TWN; i_ += G::NT) twxs_l[i_] = (TWSRC)[i_]; \
    __syncthreads();                                                     \
    (void)nyq; (void)sc;

#define HW_YK_PROLOGUE                                                   \
    a.shift((int)blockIdx.z);                                            \
    HW_YK_PROLOGUE_TW(a.twy)
#define HW_YK_PROLOGUE_TW(TWSRC)                                         \
    using C = typename ctype<T>::type;                                   \
    using G = YGeo<T, NY>;                                               \
    __shared__ C lds[G::LDS_ELEMS];                                      \
    __shared__ C twys_l[G::TWL ? G::TWN : 1];                            \
    const C* twys = G::TWL ? twys_l : (TWSRC);                           \
    const int tid = threadIdx.x;                                         \
    const int col0 = blockIdx.x * G::CC;                                 \
    if (G::TWL)                                                          \
        for (int i_ = tid; i_ < G::TWN; i_ += G::NT) twys_l[i_] = (TWSRC)[i_]; \
    __syncthreads();

// W slots used by the 2DB pipeline
enum : int {
    S_ETA = 0, S_U = 1, S_AMU = 2, S_APU = 3, S_V = 4, S_AMV = 5, S_APV = 6,
    S_FGMU = 7, S_FGPU = 8, S_FGAU = 9, S_FGMV = 10, S_FGPV = 11, S_FGAV = 12,
};

template <typename T, int NX, int NY>
__global__ void __launch_bounds__((YGeo<T, NY>::NT)) __attribute__((amdgpu_waves_per_eu(YGeo<T, NY>::WPE))) k2b_y1(Rhs2DArgs<T> a) {
    HW_YK_PROLOGUE
    // all inputs and multipliers are loaded up front; the transforms follow
    C e0[G::PE], uu[G::PE], vv[G::PE], w[G::PE];
    T cm[G::PE], cp[G::PE];
    y1_side_jobs<T, G::NT>(a, tid);
    y_load<T, NX, NY>(e0, a.eta_hat, nullptr, col0, tid);
    y_load<T, NX, NY>(uu, a.u_hat, nullptr, col0, tid);
    y_load<T, NX, NY>(vv, a.v_hat, nullptr, col0, tid);
#pragma unroll
    for (int e = 0; e < G::PE; ++e) {
        int kx, ky; size_t gi;
        YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
        const bool in = kx < NX / 2 + 1;
        cm[e] = in ? a.csq_m[gi] : (T)0;
        cp[e] = in ? a.csq_p[gi] : (T)0;
    }
    y_inv_regs<T, NX, NY>(e0, lds, twys, col0, tid);
    y_store<T, NX, NY>(a.W[S_ETA], e0, col0, tid);
    const int slot[2][3] = {{S_U, S_AMU, S_APU}, {S_V, S_AMV, S_APV}};
    #pragma unroll
    for (int f = 0; f < 2; ++f) {
        auto& v = f == 0 ? uu : vv;
#pragma unroll
        for (int e = 0; e < G::PE; ++e) w[e] = v[e];
        y_inv_regs<T, NX, NY>(w, lds, twys, col0, tid);
        y_store<T, NX, NY>(a.W[slot[f][0]], w, col0, tid);
        #pragma unroll
        for (int s = 0; s < 2; ++s) {
#pragma unroll
            for (int e = 0; e < G::PE; ++e) w[e] = cscale(v[e], s == 0 ? cm[e] : cp[e]);
            y_inv_regs<T, NX, NY>(w, lds, twys, col0, tid);
            y_store<T, NX, NY>(a.W[slot[f][1 + s]], w, col0, tid);
        }
    }
}

template <typename T, int NX, int NY>
__global__ void HW_XK_BOUNDS k2b_x1(Rhs2DArgs<T> a) {
    HW_XK_PROLOGUE
    // eta, u, v and the HSS inverses of u^ and v^.  One spectral buffer rolls
    // through the seven inputs (each inverse refills it with the next one).
    C f[G::PP], u[G::PP], gm[G::PP], gp[G::PP];
    XSpec<T, NX> b;
    x_spec_load<T, NX, NY>(a.W[S_ETA], y0, b, tid);
    ph_load<T, NX, NY>(gm, a.gam_m, y0, tid);
    ph_load<T, NX, NY>(gp, a.gam_p, y0, tid);
    x_inv_roll<T, NX, NY>(b, a.W[S_U], y0, lds, twxs, f, sc, nullptr, tid);
    ph_store<T, NX, NY>(a.eta_p, f, y0, tid);
    #pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int s_am = c == 0 ? S_AMU : S_AMV, s_ap = c == 0 ? S_APU : S_APV;
        // velocity: store it, F(g- vel), F(g+ vel)
        // (the next input is loaded after the two forward transforms: holding
        // it through them would exceed the register budget)
        x_inv_roll<T, NX, NY>(b, nullptr, y0, lds, twxs, u, sc, nullptr, tid);
        ph_store<T, NX, NY>(c == 0 ? a.u_p : a.v_p, u, y0, tid);
#pragma unroll
        for (int q = 0; q < G::PP; ++q) f[q] = pmul(gm[q], u[q]);
        x_fwd_store<T, NX, NY>(f, a.W[c == 0 ? S_FGMU : S_FGMV], y0, lds, twxs, tid);
#pragma unroll
        for (int q = 0; q < G::PP; ++q) f[q] = pmul(gp[q], u[q]);
        x_spec_load<T, NX, NY>(a.W[s_am], y0, b, tid);
        x_fwd_store<T, NX, NY>(f, a.W[c == 0 ? S_FGPU : S_FGPV], y0, lds, twxs, tid);
        // F(g- F^-1(C2- vel^) + g+ F^-1(C2+ vel^)), accumulated in u
        x_inv_roll<T, NX, NY>(b, a.W[s_ap], y0, lds, twxs, f, sc, nullptr, tid);
#pragma unroll
        for (int q = 0; q < G::PP; ++q) u[q] = pmul(gm[q], f[q]);
        x_inv_roll<T, NX, NY>(b, c == 0 ? a.W[S_V] : nullptr, y0, lds, twxs, f, sc, nullptr, tid);
#pragma unroll
        for (int q = 0; q < G::PP; ++q) u[q] = pfma(gp[q], f[q], u[q]);
        x_fwd_store<T, NX, NY>(u, a.W[c == 0 ? S_FGAU : S_FGAV], y0, lds, twxs, tid);
    }
}

// W slots after Y2
enum : int { S_M0U = 0, S_M0V = 1 };

template <typename T, int NX, int NY>
__global__ void __launch_bounds__((YGeo<T, NY>::NT)) __attribute__((amdgpu_waves_per_eu(YGeo<T, NY>::WPE))) k2b_y2(Rhs2DArgs<T> a) {
    HW_YK_PROLOGUE
    C A0[G::PE], B0[G::PE], C0[G::PE], A1[G::PE], B1[G::PE], C1[G::PE], dsum[G::PE];
    T cm[G::PE], cp[G::PE], kk0[G::PE], kk1[G::PE];
    const T ig = (T)1 / a.g;
    y_load<T, NX, NY>(A0, a.W[S_FGMU], nullptr, col0, tid);
    y_load<T, NX, NY>(B0, a.W[S_FGPU], nullptr, col0, tid);
    y_load<T, NX, NY>(C0, a.W[S_FGAU], nullptr, col0, tid);
    y_load<T, NX, NY>(A1, a.W[S_FGMV], nullptr, col0, tid);
    y_load<T, NX, NY>(B1, a.W[S_FGPV], nullptr, col0, tid);
    y_load<T, NX, NY>(C1, a.W[S_FGAV], nullptr, col0, tid);
#pragma unroll
    for (int e = 0; e < G::PE; ++e) {
        int kx, ky; size_t gi;
        YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
        const bool in = kx < NX / 2 + 1;
        cm[e] = in ? a.csq_m[gi] : (T)0;
        cp[e] = in ? a.csq_p[gi] : (T)0;
        kk0[e] = in ? a.kxh[kx] : (T)0;
        kk1[e] = in ? a.ky[ky] : (T)0;
    }
    #pragma unroll
    for (int c = 0; c < 2; ++c) {
        auto& A = c == 0 ? A0 : A1;
        auto& B = c == 0 ? B0 : B1;
        auto& Cc = c == 0 ? C0 : C1;
        y_fwd_regs<T, NX, NY>(A, lds, twys, tid);
        y_fwd_regs<T, NX, NY>(B, lds, twys, tid);
        y_fwd_regs<T, NX, NY>(Cc, lds, twys, tid);
#pragma unroll
        for (int e = 0; e < G::PE; ++e) {
            C hs = cscale(cadd(Cc[e], cadd(cscale(A[e], cm[e]), cscale(B[e], cp[e]))), (T)0.5);
            C d = cmul_i<-1>(cscale(hs, (c == 0 ? kk0[e] : kk1[e]) * ig));     // -i k HSS / g
            A[e] = d;
            dsum[e] = c == 0 ? d : cadd(dsum[e], d);
        }
        y_inv_regs<T, NX, NY>(A, lds, twys, col0, tid);
        y_store<T, NX, NY>(a.W[c == 0 ? S_M0U : S_M0V], A, col0, tid);
    }
    y_store<T, NX, NY>(a.dH1, dsum, col0, tid);
}

// W slots after X3
enum : int { S_P1 = 2, S_Q1 = 3, S_G1M = 4, S_G1P = 5, S_P2 = 6, S_Q2 = 7, S_G2M = 8, S_G2P = 9, S_E = 10 };

template <typename T, int NX, int NY, bool H3 = false>
__global__ void HW_XK_BOUNDS k2b_x3(Rhs2DArgs<T> a) {
    HW_XK_PROLOGUE
    // u and v are read twice (for E and for the products) so that at most
    // four pair arrays are live during a transform
    C m0u[G::PP], m0v[G::PP], w[G::PP], x[G::PP];
    XSpec<T, NX> b;
    x_spec_load<T, NX, NY>(a.W[S_M0U], y0, b, tid);
    x_inv_roll<T, NX, NY>(b, a.W[S_M0V], y0, lds, twxs, m0u, sc, nullptr, tid);
    if constexpr (H3) ph_store<T, NX, NY>(a.m0u_p, m0u, y0, tid);
    ph_load<T, NX, NY>(w, a.u_p, y0, tid);
    x_inv_roll<T, NX, NY>(b, nullptr, y0, lds, twxs, m0v, sc, nullptr, tid);
    if constexpr (H3) ph_store<T, NX, NY>(a.m0v_p, m0v, y0, tid);
    // E = 0.5 (u^2 + v^2 - Lphi^2) w, Lphi = M0u + M0v  (2d/fun_odesolver.c:3426)
    ph_load<T, NX, NY>(x, a.v_p, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) {
        C lp = cadd(m0u[q], m0v[q]);
        x[q] = cmk<C>((w[q].x * w[q].x + x[q].x * x[q].x) - lp.x * lp.x, (w[q].y * w[q].y + x[q].y * x[q].y) - lp.y * lp.y);
    }
    ph_load<T, NX, NY>(w, a.wallchar, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) x[q] = cmk<C>((T)0.5 * x[q].x * w[q].x, (T)0.5 * x[q].y * w[q].y);
    x_fwd_store<T, NX, NY>(x, a.W[S_E], y0, lds, twxs, tid);
    // w <- eta w; Q = F(eta u w), F(eta v w)
    ph_load<T, NX, NY>(x, a.eta_p, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) w[q] = pmul(x[q], w[q]);
    ph_load<T, NX, NY>(x, a.u_p, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) x[q] = pmul(w[q], x[q]);
    x_fwd_store<T, NX, NY>(x, a.W[S_Q1], y0, lds, twxs, tid);
    ph_load<T, NX, NY>(x, a.v_p, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) x[q] = pmul(w[q], x[q]);
    x_fwd_store<T, NX, NY>(x, a.W[S_Q2], y0, lds, twxs, tid);
    // P = F(eta M0 w), G = F(g-+ eta M0 w)
#pragma unroll
    for (int q = 0; q < G::PP; ++q) { m0u[q] = pmul(w[q], m0u[q]); m0v[q] = pmul(w[q], m0v[q]); }
    x_fwd_store<T, NX, NY>(m0u, a.W[S_P1], y0, lds, twxs, tid);
    x_fwd_store<T, NX, NY>(m0v, a.W[S_P2], y0, lds, twxs, tid);
    #pragma unroll
    for (int s = 0; s < 2; ++s) {
        ph_load<T, NX, NY>(w, s == 0 ? a.gam_m : a.gam_p, y0, tid);
#pragma unroll
        for (int q = 0; q < G::PP; ++q) x[q] = pmul(w[q], m0u[q]);
        x_fwd_store<T, NX, NY>(x, a.W[s == 0 ? S_G1M : S_G1P], y0, lds, twxs, tid);
#pragma unroll
        for (int q = 0; q < G::PP; ++q) x[q] = pmul(w[q], m0v[q]);
        x_fwd_store<T, NX, NY>(x, a.W[s == 0 ? S_G2M : S_G2P], y0, lds, twxs, tid);
    }
}

// W slots after Y4 (inputs S_P1.. are consumed)
enum : int { S_B1M = 11, S_B1P = 12, S_B2M = 13, S_B2P = 14, S_AX = 0, S_AY = 1 };

template <typename T, int NX, int NY>
__global__ void __launch_bounds__((YGeo<T, NY>::NT)) __attribute__((amdgpu_waves_per_eu(YGeo<T, NY>::WPE))) k2b_y4(Rhs2DArgs<T> a) {
    HW_YK_PROLOGUE
    // Every field's loads are issued one transform ahead of its use (the
    // arrays rotate roles).
    C A[G::PE], B[G::PE], D[G::PE], F[G::PE], part[G::PE];
    const T ig = (T)1 / a.g;
    // Om2-+ at element e (read where used: holding them costs more registers than the cached re-reads)
    auto om = [&](int s, int e) -> T {
        int kx, ky; size_t gi;
        YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
        return kx < NX / 2 + 1 ? (s == 0 ? a.om2_m[gi] : a.om2_p[gi]) : (T)0;
    };
    // -i k (eta u)^ - (Om2- F(g- p) + Om2+ F(g+ p)) / (2 g), k = kx (c = 0) or ky (c = 1)
    auto combine = [&](int c, const C* Qv, const C* Gm, const C* Gp) {
#pragma unroll
        for (int e = 0; e < G::PE; ++e) {
            int kx, ky; size_t gi;
            YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
            if (kx < NX / 2 + 1) {
                T kk = c == 0 ? a.kxh[kx] : a.ky[ky];
                C t1 = cmul_i<-1>(cscale(Qv[e], kk));
                C t2 = cscale(cadd(cscale(Gm[e], om(0, e)), cscale(Gp[e], om(1, e))), (T)0.5 * ig);
                C d = csub(t1, t2);
                part[e] = c == 0 ? d : cadd(part[e], d);
            } else if (c == 0) {
                part[e] = cmk<C>((T)0, (T)0);
            }
        }
    };
    // inverse y of Om2-+ F(eta M0 w): w uses the scratch array
    auto om_inverses = [&](const C* Pv, C* w, int s_m, int s_p) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
#pragma unroll
            for (int e = 0; e < G::PE; ++e) w[e] = cscale(Pv[e], om(s, e));
            y_inv_regs<T, NX, NY>(*reinterpret_cast<C(*)[G::PE]>(w), lds, twys, col0, tid);
            y_store<T, NX, NY>(a.W[s == 0 ? s_m : s_p], *reinterpret_cast<C(*)[G::PE]>(w), col0, tid);
        }
    };
    // x momentum terms: A = P1, B = Q1, D = G1M, F = G1P
    y_load<T, NX, NY>(A, a.W[S_P1], nullptr, col0, tid);
    y_load<T, NX, NY>(B, a.W[S_Q1], nullptr, col0, tid);
    y_fwd_regs<T, NX, NY>(A, lds, twys, tid);
    y_load<T, NX, NY>(D, a.W[S_G1M], nullptr, col0, tid);
    y_fwd_regs<T, NX, NY>(B, lds, twys, tid);
    y_load<T, NX, NY>(F, a.W[S_G1P], nullptr, col0, tid);
    y_fwd_regs<T, NX, NY>(D, lds, twys, tid);
    y_fwd_regs<T, NX, NY>(F, lds, twys, tid);
    combine(0, B, D, F);
    y_load<T, NX, NY>(B, a.W[S_Q2], nullptr, col0, tid);
    y_load<T, NX, NY>(D, a.W[S_P2], nullptr, col0, tid);
    om_inverses(A, F, S_B1M, S_B1P);
    // y momentum terms: D = P2, B = Q2, A = G2M, F = G2P
    y_fwd_regs<T, NX, NY>(D, lds, twys, tid);
    y_load<T, NX, NY>(A, a.W[S_G2M], nullptr, col0, tid);
    y_fwd_regs<T, NX, NY>(B, lds, twys, tid);
    y_load<T, NX, NY>(F, a.W[S_G2P], nullptr, col0, tid);
    y_fwd_regs<T, NX, NY>(A, lds, twys, tid);
    y_fwd_regs<T, NX, NY>(F, lds, twys, tid);
    combine(1, B, A, F);
    y_load<T, NX, NY>(B, a.W[S_E], nullptr, col0, tid);
    om_inverses(D, F, S_B2M, S_B2P);
    y_store<T, NX, NY>(a.dH2p, part, col0, tid);
    // adjust inputs i kx E^, i ky E^
    y_fwd_regs<T, NX, NY>(B, lds, twys, tid);
    #pragma unroll
    for (int s = 0; s < 2; ++s) {
#pragma unroll
        for (int e = 0; e < G::PE; ++e) {
            int kx, ky; size_t gi;
            YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
            T kk = (kx < NX / 2 + 1) ? (s == 0 ? a.kxh[kx] : a.ky[ky]) : (T)0;
            F[e] = cmul_i<1>(cscale(B[e], kk));
        }
        y_inv_regs<T, NX, NY>(F, lds, twys, col0, tid);
        y_store<T, NX, NY>(a.W[s == 0 ? S_AX : S_AY], F, col0, tid);
    }
}

// W slots after X5
enum : int { S_ACC1 = 2, S_ACC2 = 3, S_ZU = 4, S_ZV = 5 };


// output transform of a velocity equation: f1 = Sb - chi A - fbdy7 vel,
// f2 = -chi (-1)^x s(y)   (the imaginary part of fun_adjust's inverse).
// b holds the adjust input A^; it is refilled with `next`.
template <typename T, int NX, int NY>
__device__ __forceinline__ void x_vel_out(const Rhs2DArgs<T>& a, XSpec<T, NX>& b, const typename ctype<T>::type* next,
                                          const T* vel, const T* sb, int so, int y0, typename ctype<T>::type* lds,
                                          const typename ctype<T>::type* twl, T* nyq, int tid) {
    using C = typename ctype<T>::type;
    using G = XGeo<T, NX>;
    const T sc = (T)1 / ((T)NX * (T)NY);
    C f1[G::PP], tc[G::PP], t[G::PP];
    x_inv_roll<T, NX, NY>(b, next, y0, lds, twl, f1, sc, nyq, tid);
    ph_load<T, NX, NY>(tc, a.chi, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) f1[q] = cmk<C>(-tc[q].x * f1[q].x, -tc[q].y * f1[q].y);
    if (sb) {
        ph_load<T, NX, NY>(t, sb, y0, tid);
#pragma unroll
        for (int q = 0; q < G::PP; ++q) f1[q] = cadd(f1[q], t[q]);
    }
    {
        C v[G::PP];
        ph_load<T, NX, NY>(t, a.fbdy7, y0, tid);
        ph_load<T, NX, NY>(v, vel, y0, tid);
#pragma unroll
        for (int q = 0; q < G::PP; ++q) f1[q] = csub(f1[q], pmul(t[q], v[q]));
    }
    x_fwd_adj_store<T, NX, NY>(f1, tc, (T)-1, nyq, a.W[so], y0, lds, twl, tid);
}

template <typename T, int NX, int NY>
__global__ void HW_XK_BOUNDS k2b_x5(Rhs2DArgs<T> a) {
    HW_XK_PROLOGUE
    // F(g- b- + g+ b+) of both HSS tails, then the two velocity output
    // transforms; one spectral buffer rolls through the six inputs
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
    x_inv_roll<T, NX, NY>(b, a.W[S_AX], y0, lds, twxs, f, sc, nullptr, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) acc[q] = pfma(gp[q], f[q], acc[q]);
    x_fwd_store<T, NX, NY>(acc, a.W[S_ACC2], y0, lds, twxs, tid);
    x_vel_out<T, NX, NY>(a, b, a.W[S_AY], a.u_p, a.sbx, S_ZU, y0, lds, twxs, nyq, tid);
    x_vel_out<T, NX, NY>(a, b, nullptr, a.v_p, a.sby, S_ZV, y0, lds, twxs, nyq, tid);
}

// W slots after Y6
enum : int { S_AETA = 6 };

// HS3 slots (two-depth): the HS2 tail of the y momentum moves to slot 8 so
// that F(eta u w), F(eta v w) stay readable, and the HS3 tail is slot 14.
enum : int { S3_ACC2 = 8, S3_ACC3 = 14 };

template <typename T, int NX, int NY, bool H3 = false>
__global__ void __launch_bounds__((YGeo<T, NY>::NT)) __attribute__((amdgpu_waves_per_eu(YGeo<T, NY>::WPE))) k2b_y6(Rhs2DArgs<T> a) {
    HW_YK_PROLOGUE
    C F1[G::PE], F2[G::PE], Z[G::PE], eh[G::PE], Zu[G::PE], Zv[G::PE];
    T al[G::PE], kk0[G::PE], kk1[G::PE];
    const T ig = (T)1 / a.g;
    y_load<T, NX, NY>(F1, a.W[S_ACC1], nullptr, col0, tid);
    y_load<T, NX, NY>(F2, a.W[H3 ? (int)S3_ACC2 : (int)S_ACC2], nullptr, col0, tid);
    y_load<T, NX, NY>(Z, a.dH2p, nullptr, col0, tid);
    y_load<T, NX, NY>(eh, a.eta_hat, nullptr, col0, tid);
    y_load<T, NX, NY>(Zu, a.W[S_ZU], nullptr, col0, tid);
    y_load<T, NX, NY>(Zv, a.W[S_ZV], nullptr, col0, tid);
#pragma unroll
    for (int e = 0; e < G::PE; ++e) {
        int kx, ky; size_t gi;
        YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
        const bool in = kx < NX / 2 + 1;
        al[e] = in ? a.aal[gi] : (T)0;
        kk0[e] = in ? a.kxh[kx] : (T)0;
        kk1[e] = in ? a.ky[ky] : (T)0;
    }
    y_fwd_regs<T, NX, NY>(F1, lds, twys, tid);
    y_fwd_regs<T, NX, NY>(F2, lds, twys, tid);
    if constexpr (H3) {
        // the HS3 tail F(g- b3- + g+ b3+) enters with the opposite sign
#pragma unroll
        for (int e = 0; e < G::PE; ++e) F1[e] = cadd(F1[e], F2[e]);
        y_load<T, NX, NY>(F2, a.W[S3_ACC3], nullptr, col0, tid);
        y_fwd_regs<T, NX, NY>(F2, lds, twys, tid);
#pragma unroll
        for (int e = 0; e < G::PE; ++e) F2[e] = cmk<C>(-F2[e].x, -F2[e].y);
    }
#pragma unroll
    for (int e = 0; e < G::PE; ++e) Z[e] = csub(Z[e], cscale(cadd(F1[e], F2[e]), (T)0.5 * ig));
    y_inv_regs<T, NX, NY>(Z, lds, twys, col0, tid);
    y_store<T, NX, NY>(a.W[S_AETA], Z, col0, tid);
    #pragma unroll
    for (int c = 0; c < 2; ++c) {
        auto& Zc = c == 0 ? Zu : Zv;
        y_fwd_regs<T, NX, NY>(Zc, lds, twys, tid);
#pragma unroll
        for (int e = 0; e < G::PE; ++e) {
            C lin = cmul_i<-1>(cscale(eh[e], a.g * (c == 0 ? kk0[e] : kk1[e])));
            Zc[e] = cscale(cadd(lin, Zc[e]), al[e]);
        }
        y_store<T, NX, NY>(c == 0 ? a.k_u : a.k_v, Zc, col0, tid);
    }
}

// source term of fun_source (2d/fun_odesolver.c:909-960) for the pair q of this thread
template <typename T, int NX>
__device__ __forceinline__ void src_pair(const Rhs2DArgs<T>& a, int y0, int pos, typename ctype<T>::type ng,
                                         typename ctype<T>::type sg, T& s0, T& s1) {
    using G = XGeo<T, NX>;
    int r = pos / G::L, n = pos % G::L;
    T sig0, sig1, sk0, sk1;
    if (a.vertical) {
        sig0 = sig1 = a.src_sig[y0 + r];
        sk0 = sk1 = a.src_skew[y0 + r];
    } else {
        sig0 = a.src_sig[2 * n]; sig1 = a.src_sig[2 * n + 1];
        sk0 = a.src_skew[2 * n]; sk1 = a.src_skew[2 * n + 1];
    }
    if (a.propagation == 2) {
        s0 = 2 * sig0 * ng.x; s1 = 2 * sig1 * ng.y;
    } else {
        T sgn = a.propagation == 0 ? (T)-1 : (T)1;
        s0 = sig0 * ng.x + sgn * sk0 * sg.x;
        s1 = sig1 * ng.y + sgn * sk1 * sg.y;
    }
}

// output transform of the eta equation: f1 = S(t) - fbdy7 eta + chi A,
// f2 = chi (-1)^x s(y).  b holds A^; it is refilled with `next`.
template <typename T, int NX, int NY>
__device__ __forceinline__ void x_eta_out(const Rhs2DArgs<T>& a, XSpec<T, NX>& b, const typename ctype<T>::type* next,
                                          int so, int y0, typename ctype<T>::type* lds,
                                          const typename ctype<T>::type* twl, T* nyq, int tid) {
    using C = typename ctype<T>::type;
    using G = XGeo<T, NX>;
    const T sc = (T)1 / ((T)NX * (T)NY);
    C f1[G::PP], tc[G::PP], t[G::PP], e[G::PP];
    x_inv_roll<T, NX, NY>(b, next, y0, lds, twl, f1, sc, nyq, tid);
    ph_load<T, NX, NY>(tc, a.chi, y0, tid);
    ph_load<T, NX, NY>(t, a.fbdy7, y0, tid);
    ph_load<T, NX, NY>(e, a.eta_p, y0, tid);
    if (a.s_eta_out && tid < G::RR) a.s_eta_out[y0 + tid] = nyq[tid];
#pragma unroll
    for (int q = 0; q < G::PP; ++q) f1[q] = csub(pmul(tc[q], f1[q]), pmul(t[q], e[q]));
    if (a.src_on) {
        ph_load<T, NX, NY>(t, a.new_gam, y0, tid);
        ph_load<T, NX, NY>(e, a.skew_gam, y0, tid);
#pragma unroll
        for (int q = 0; q < G::PP; ++q) {
            T s0, s1;
            src_pair<T, NX>(a, y0, G::ppos(tid, q), t[q], e[q], s0, s1);
            f1[q] = cadd(f1[q], cmk<C>(s0, s1));
        }
    }
    x_fwd_adj_store<T, NX, NY>(f1, tc, (T)1, nyq, a.W[so], y0, lds, twl, tid);
}

template <typename T, int NX, int NY>
__global__ void HW_XK_BOUNDS k2b_x7(Rhs2DArgs<T> a) {
    HW_XK_PROLOGUE
    XSpec<T, NX> b;
    x_spec_load<T, NX, NY>(a.W[S_AETA], y0, b, tid);
    x_eta_out<T, NX, NY>(a, b, nullptr, S_ZU, y0, lds, twxs, nyq, tid);
}

template <typename T, int NX, int NY>
__global__ void __launch_bounds__((YGeo<T, NY>::NT)) k2b_y8(Rhs2DArgs<T> a) {
    HW_YK_PROLOGUE
    C Z[G::PE], d[G::PE];
    y_load<T, NX, NY>(Z, a.W[S_ZU], nullptr, col0, tid);
    y_fwd_regs<T, NX, NY>(Z, lds, twys, tid);
    y_load<T, NX, NY>(d, a.dH1, nullptr, col0, tid);
#pragma unroll
    for (int e = 0; e < G::PE; ++e) {
        int kx, ky; size_t gi;
        YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
        if (kx < NX / 2 + 1) {
            C pre = cadd(d[e], Z[e]);
            if (a.eta_col0 && kx == 0) a.eta_col0[ky] = pre;
            Z[e] = pre;
        }
    }
    y_eta_mask<T, NX, NY>(Z, a, lds, col0, tid);
    y_store<T, NX, NY>(a.k_eta, Z, col0, tid);
}

// ========================================================== 2DF kernels ==
// rhs_AB2_2DF_uv without walls: delphiH1 = -i C2d^2/g (kx u^ + ky v^) is
// pointwise on the input; M0u^ = -i kx C2d^2/g u^, M0v^ likewise
// (2d/fun_odesolver.c:2223-2236, 3172-3178).
enum : int { F_ETA = 0, F_U = 1, F_V = 2, F_M0U = 3, F_M0V = 4 };

template <typename T, int NX, int NY>
__global__ void __launch_bounds__((YGeo<T, NY>::NT)) __attribute__((amdgpu_waves_per_eu(YGeo<T, NY>::WPE))) k2f_y1(Rhs2DArgs<T> a) {
    HW_YK_PROLOGUE
    C v[G::PE], w[G::PE];
    y1_side_jobs<T, G::NT>(a, tid);
    y_load<T, NX, NY>(v, a.eta_hat, nullptr, col0, tid);
    y_inv_regs<T, NX, NY>(v, lds, twys, col0, tid);
    y_store<T, NX, NY>(a.W[F_ETA], v, col0, tid);
    #pragma unroll
    for (int c = 0; c < 2; ++c) {
        y_load<T, NX, NY>(v, c == 0 ? a.u_hat : a.v_hat, nullptr, col0, tid);
#pragma unroll
        for (int e = 0; e < G::PE; ++e) {
            int kx, ky; size_t gi;
            YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
            T kk = (kx < NX / 2 + 1) ? (c == 0 ? a.kxh[kx] : a.ky[ky]) * a.c2d2g[gi] : (T)0;
            w[e] = cmul_i<-1>(cscale(v[e], kk));
        }
        y_inv_regs<T, NX, NY>(v, lds, twys, col0, tid);
        y_store<T, NX, NY>(a.W[c == 0 ? F_U : F_V], v, col0, tid);
        y_inv_regs<T, NX, NY>(w, lds, twys, col0, tid);
        y_store<T, NX, NY>(a.W[c == 0 ? F_M0U : F_M0V], w, col0, tid);
    }
}

enum : int { F_P1 = 5, F_Q1 = 6, F_P2 = 7, F_Q2 = 8, F_E = 9 };

template <typename T, int NX, int NY, bool H3 = false>
__global__ void HW_XK_BOUNDS k2f_x2(Rhs2DArgs<T> a) {
    HW_XK_PROLOGUE
    C ew[G::PP], m[G::PP], q2[G::PP], lp[G::PP];
    XSpec<T, NX> b;
    x_spec_load<T, NX, NY>(a.W[F_ETA], y0, b, tid);
    x_inv_roll<T, NX, NY>(b, a.W[F_U], y0, lds, twxs, ew, sc, nullptr, tid);
    ph_store<T, NX, NY>(a.eta_p, ew, y0, tid);
    ph_load<T, NX, NY>(m, a.wallchar, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) ew[q] = pmul(ew[q], m[q]);   // eta w
    x_inv_roll<T, NX, NY>(b, a.W[F_V], y0, lds, twxs, m, sc, nullptr, tid);
    ph_store<T, NX, NY>(a.u_p, m, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) { q2[q] = pmul(m[q], m[q]); m[q] = pmul(ew[q], m[q]); }
    x_fwd_store<T, NX, NY>(m, a.W[F_Q1], y0, lds, twxs, tid);
    x_inv_roll<T, NX, NY>(b, a.W[F_M0U], y0, lds, twxs, m, sc, nullptr, tid);
    ph_store<T, NX, NY>(a.v_p, m, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) { q2[q] = pfma(m[q], m[q], q2[q]); m[q] = pmul(ew[q], m[q]); }   // u^2 + v^2
    x_fwd_store<T, NX, NY>(m, a.W[F_Q2], y0, lds, twxs, tid);
    x_inv_roll<T, NX, NY>(b, a.W[F_M0V], y0, lds, twxs, lp, sc, nullptr, tid);
    if constexpr (H3) ph_store<T, NX, NY>(a.m0u_p, lp, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) m[q] = pmul(ew[q], lp[q]);
    x_fwd_store<T, NX, NY>(m, a.W[F_P1], y0, lds, twxs, tid);
    x_inv_roll<T, NX, NY>(b, nullptr, y0, lds, twxs, m, sc, nullptr, tid);
    if constexpr (H3) ph_store<T, NX, NY>(a.m0v_p, m, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q) { lp[q] = cadd(lp[q], m[q]); m[q] = pmul(ew[q], m[q]); }
    x_fwd_store<T, NX, NY>(m, a.W[F_P2], y0, lds, twxs, tid);
    ph_load<T, NX, NY>(m, a.wallchar, y0, tid);
#pragma unroll
    for (int q = 0; q < G::PP; ++q)
        m[q] = cmk<C>((T)0.5 * (q2[q].x - lp[q].x * lp[q].x) * m[q].x, (T)0.5 * (q2[q].y - lp[q].y * lp[q].y) * m[q].y);
    x_fwd_store<T, NX, NY>(m, a.W[F_E], y0, lds, twxs, tid);
}

enum : int { F_AETA = 0, F_AX = 1, F_AY = 2 };

template <typename T, int NX, int NY>
__global__ void __launch_bounds__((YGeo<T, NY>::NT)) __attribute__((amdgpu_waves_per_eu(YGeo<T, NY>::WPE))) k2f_y3(Rhs2DArgs<T> a) {
    HW_YK_PROLOGUE
    C P[G::PE], Q[G::PE], acc[G::PE];
    const int sl[2][2] = {{F_P1, F_Q1}, {F_P2, F_Q2}};
    #pragma unroll
    for (int c = 0; c < 2; ++c) {
        y_load<T, NX, NY>(P, a.W[sl[c][0]], nullptr, col0, tid);
        y_fwd_regs<T, NX, NY>(P, lds, twys, tid);
        y_load<T, NX, NY>(Q, a.W[sl[c][1]], nullptr, col0, tid);
        y_fwd_regs<T, NX, NY>(Q, lds, twys, tid);
#pragma unroll
        for (int e = 0; e < G::PE; ++e) {
            int kx, ky; size_t gi;
            YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
            if (kx < NX / 2 + 1) {
                T kk = c == 0 ? a.kxh[kx] : a.ky[ky];
                // -L2d (eta M0)^ - i k (eta u)^   (2d/fun_odesolver.c:3293-3297)
                C d = csub(cmul_i<-1>(cscale(Q[e], kk)), cscale(P[e], a.l2d[gi]));
                acc[e] = c == 0 ? d : cadd(acc[e], d);
            } else {
                acc[e] = cmk<C>((T)0, (T)0);
            }
        }
    }
    y_inv_regs<T, NX, NY>(acc, lds, twys, col0, tid);
    y_store<T, NX, NY>(a.W[F_AETA], acc, col0, tid);
    y_load<T, NX, NY>(P, a.W[F_E], nullptr, col0, tid);
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
        y_store<T, NX, NY>(a.W[s == 0 ? F_AX : F_AY], Q, col0, tid);
    }
}

enum : int { F_ZETA = 3, F_ZU = 4, F_ZV = 5 };

template <typename T, int NX, int NY>
__global__ void HW_XK_BOUNDS k2f_x4(Rhs2DArgs<T> a) {
    HW_XK_PROLOGUE
    XSpec<T, NX> b;
    x_spec_load<T, NX, NY>(a.W[F_AETA], y0, b, tid);
    x_eta_out<T, NX, NY>(a, b, a.W[F_AX], F_ZETA, y0, lds, twxs, nyq, tid);
    x_vel_out<T, NX, NY>(a, b, a.W[F_AY], a.u_p, a.sbx, F_ZU, y0, lds, twxs, nyq, tid);
    x_vel_out<T, NX, NY>(a, b, nullptr, a.v_p, a.sby, F_ZV, y0, lds, twxs, nyq, tid);
}

template <typename T, int NX, int NY>
__global__ void __launch_bounds__((YGeo<T, NY>::NT)) __attribute__((amdgpu_waves_per_eu(YGeo<T, NY>::WPE))) k2f_y5(Rhs2DArgs<T> a) {
    HW_YK_PROLOGUE
    C Z[G::PE], eh[G::PE], uh[G::PE], vh[G::PE];
    y_load<T, NX, NY>(eh, a.eta_hat, nullptr, col0, tid);
    y_load<T, NX, NY>(uh, a.u_hat, nullptr, col0, tid);
    y_load<T, NX, NY>(vh, a.v_hat, nullptr, col0, tid);
    const int zs[3] = {F_ZETA, F_ZU, F_ZV};
    C* outs[3] = {a.k_eta, a.k_u, a.k_v};
    #pragma unroll
    for (int c = 0; c < 3; ++c) {
        y_load<T, NX, NY>(Z, a.W[zs[c]], nullptr, col0, tid);
        y_fwd_regs<T, NX, NY>(Z, lds, twys, tid);
#pragma unroll
        for (int e = 0; e < G::PE; ++e) {
            int kx, ky; size_t gi;
            YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
            if (kx < NX / 2 + 1) {
                C lin;
                if (c == 0) {
                    // delphiH1 = -i C2d^2/g (kx u^ + ky v^)
                    C s = cadd(cscale(uh[e], a.kxh[kx]), cscale(vh[e], a.ky[ky]));
                    lin = cmul_i<-1>(cscale(s, a.c2d2g[gi]));
                } else {
                    T kk = c == 1 ? a.kxh[kx] : a.ky[ky];
                    lin = cmul_i<-1>(cscale(eh[e], a.g * kk));
                }
                C pre = cadd(lin, Z[e]);
                if (c == 0 && a.eta_col0 && kx == 0) a.eta_col0[ky] = pre;
                Z[e] = c == 0 ? pre : cscale(pre, a.aal[gi]);
            }
        }
        if (c == 0) y_eta_mask<T, NX, NY>(Z, a, lds, col0, tid);
        y_store<T, NX, NY>(outs[c], Z, col0, tid);
    }
}

// ========================================================== HS1 kernels ==
// rhs_AB1_2DB_uv / rhs_AB1_2DF_uv (2d/fun_odesolver.c:4350-4456), the linear
// model: d eta^/dt = aal (S^ + delphiH1 - F(7 Cp fbdy eta)),
// d u^/dt = aal (-i g kx eta^ - F(7 Cp fbdy u)), v likewise; no breaking.
// Two-depth: Y1, X1 and Y2 of the HS2 pipeline give delphiH1 (dH1) and the
// physical eta, u, v; flat: the Y1 of the flat pipeline gives the x-
// intermediates of eta^, u^, v^ and delphiH1 is pointwise.
enum : int { H1_ZE = 13, H1_ZU = 14, H1_ZV = 15 };

template <typename T, int NX, int NY>
__global__ void HW_XK_BOUNDS k1_x(Rhs2DArgs<T> a, int from_phys) {
    HW_XK_PROLOGUE
    // forward x of source - fbdy7 eta, -fbdy7 u, -fbdy7 v
    C f[G::PP], d[G::PP];
    ph_load<T, NX, NY>(d, a.fbdy7, y0, tid);
    // unrolled: a runtime index into a.W would move the argument struct to scratch
    #pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (from_phys) ph_load<T, NX, NY>(f, c == 0 ? a.eta_p : (c == 1 ? a.u_p : a.v_p), y0, tid);
        else x_inverse<T, NX, NY>(a.W[c == 0 ? F_ETA : (c == 1 ? F_U : F_V)], y0, lds, twxs, f, sc, nullptr, tid);
#pragma unroll
        for (int q = 0; q < G::PP; ++q) f[q] = cmk<C>(-d[q].x * f[q].x, -d[q].y * f[q].y);
        if (c == 0 && a.src_on) {
            C ng[G::PP], sg[G::PP];
            ph_load<T, NX, NY>(ng, a.new_gam, y0, tid);
            ph_load<T, NX, NY>(sg, a.skew_gam, y0, tid);
#pragma unroll
            for (int q = 0; q < G::PP; ++q) {
                T s0, s1;
                src_pair<T, NX>(a, y0, G::ppos(tid, q), ng[q], sg[q], s0, s1);
                f[q] = cadd(f[q], cmk<C>(s0, s1));
            }
        }
        x_fwd_store<T, NX, NY>(f, a.W[H1_ZE + c], y0, lds, twxs, tid);
    }
}

template <typename T, int NX, int NY>
__global__ void __launch_bounds__((YGeo<T, NY>::NT)) k1_y(Rhs2DArgs<T> a, int flat) {
    HW_YK_PROLOGUE
    C Z[G::PE], eh[G::PE];
    y_load<T, NX, NY>(eh, a.eta_hat, nullptr, col0, tid);
    C* outs[3] = {a.k_eta, a.k_u, a.k_v};
    #pragma unroll
    for (int c = 0; c < 3; ++c) {
        y_load<T, NX, NY>(Z, a.W[H1_ZE + c], nullptr, col0, tid);
        y_fwd_regs<T, NX, NY>(Z, lds, twys, tid);
#pragma unroll
        for (int e = 0; e < G::PE; ++e) {
            int kx, ky; size_t gi;
            YIdx<T, NX, NY>::get(col0, tid, e, kx, ky, gi);
            if (kx < NX / 2 + 1) {
                C lin;
                if (c == 0) {
                    if (flat) {
                        // delphiH1 = -i C2d^2/g (kx u^ + ky v^)  (2d/fun_odesolver.c:2223-2236)
                        C s = cadd(cscale(a.u_hat[gi], a.kxh[kx]), cscale(a.v_hat[gi], a.ky[ky]));
                        lin = cmul_i<-1>(cscale(s, a.c2d2g[gi]));
                    } else {
                        lin = a.dH1[gi];
                    }
                } else {
                    lin = cmul_i<-1>(cscale(eh[e], a.g * (c == 1 ? a.kxh[kx] : a.ky[ky])));
                }
                Z[e] = cscale(cadd(lin, Z[e]), a.aal[gi]);
            }
        }
        y_store<T, NX, NY>(outs[c], Z, col0, tid);
    }
}

}  // namespace hw
