// solver2d.h -- host side of the 2-D path: device buffers, the RHS launch
// sequence, the GSL-semantics RKF45 driver and boundary conversions.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "../../include/hawassi_hip.h"
#include "aux_kernels.h"
#include "breaking2d.h"
#include "crest2d.h"
#include "rhs2d_3depth.h"
#include <algorithm>
#include <climits>

namespace hw {

void set_error(const char* fmt, ...);

#define HW_TRY(expr)                                                                 \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) {                                                      \
            ::hw::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),  \
                            __FILE__, __LINE__);                                     \
            return HW_ERR_HIP;                                                       \
        }                                                                            \
    } while (0)

#define HW_LAUNCH_CHECK()                                                            \
    do {                                                                             \
        hipError_t e_ = hipGetLastError();                                           \
        if (e_ != hipSuccess) {                                                      \
            ::hw::set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), \
                            __FILE__, __LINE__);                                     \
            return HW_ERR_HIP;                                                       \
        }                                                                            \
    } while (0)

// Fehlberg tableau (GSL rkf45)
static constexpr double RK_AH[5] = {1.0 / 4.0, 3.0 / 8.0, 12.0 / 13.0, 1.0, 1.0 / 2.0};
static constexpr double RK_B3[2] = {3.0 / 32.0, 9.0 / 32.0};
static constexpr double RK_B4[3] = {1932.0 / 2197.0, -7200.0 / 2197.0, 7296.0 / 2197.0};
static constexpr double RK_B5[4] = {8341.0 / 4104.0, -32832.0 / 4104.0, 29440.0 / 4104.0, -845.0 / 4104.0};
static constexpr double RK_B6[5] = {-6080.0 / 20520.0, 41040.0 / 20520.0, -28352.0 / 20520.0, 9295.0 / 20520.0,
                                    -5643.0 / 20520.0};

inline int bsearch_interval(const std::vector<double>& x, double t) {
    int ilo = 0, ihi = (int)x.size() - 1;
    while (ihi > ilo + 1) {
        int i = (ihi + ilo) / 2;
        if (x[i] > t) ihi = i; else ilo = i;
    }
    return ilo;
}

// natural cubic spline (gsl cspline): c_i = M_i / 2 for the sample times x
inline void spline_half_second(const double* x, const double* yv, int n, double* cc) {
    for (int i = 0; i < n; ++i) cc[i] = 0.0;
    if (n < 3) return;
    int m = n - 2;
    std::vector<double> diag(m), off(m), rhs(m);
    for (int i = 0; i < m; ++i) {
        double h0 = x[i + 1] - x[i], h1 = x[i + 2] - x[i + 1];
        diag[i] = 2.0 * (h0 + h1);
        off[i] = h1;
        rhs[i] = 3.0 * ((yv[i + 2] - yv[i + 1]) / h1 - (yv[i + 1] - yv[i]) / h0);
    }
    for (int i = 1; i < m; ++i) {
        double w = off[i - 1] / diag[i - 1];
        diag[i] -= w * off[i - 1];
        rhs[i] -= w * rhs[i - 1];
    }
    cc[m] = rhs[m - 1] / diag[m - 1];
    for (int i = m - 2; i >= 0; --i) cc[i + 1] = (rhs[i] - off[i] * cc[i + 2]) / diag[i];
}

struct Solver2D {
    virtual ~Solver2D() {}
    virtual int init(const hw_config_2d* c) = 0;
    virtual int rhs_full_dev(double t, const double* d_in, double* d_out) = 0;
    virtual int set_state_dev(const double* d_full) = 0;
    virtual int get_state_dev(double* d_full) = 0;
    virtual int evolve_apply(double* t, double t1, double* h) = 0;
    // evolve_apply in pieces, for advancing several contexts in lockstep
    virtual int attempt_issue(double t0, double h0, bool need_k0) = 0;
    virtual int attempt_result(double* rmax) = 0;
    virtual int accept_step(double t) = 0;
    virtual int fields_dev(double* eta, double* u, double* v) = 0;
    virtual int breaking_state(double* B, int* flagbr) = 0;
    // device step control replayed from a HIP graph (HW_STEP_GRAPH): the
    // driver loop to t1 in pieces, for one context or several in lockstep;
    // t and h hold one value per member
    virtual int graph_begin(const double* t, double t1, const double* h) = 0;
    virtual int graph_launch() = 0;                  // a batch of attempts + the state copy-back
    virtual int graph_poll(bool* done) = 0;          // waits for it
    virtual int graph_finish(double* t, double* h) = 0;
    // member z's state of a context holding several (nb > 1)
    virtual int set_member_dev(const double* d_full, int z) = 0;
    virtual int get_member_dev(double* d_full, int z) = 0;
    int nb = 1;            // states of this model advanced together (members)
    int step_mode = 0;     // HW_STEP_HOST or HW_STEP_GRAPH
    double tprev = 0.0;    // the reference's global tprev
    int nx = 0, ny = 0, prec = 0;
    hipStream_t stream = nullptr;
    hw_stats st{};
    double* d_full_a = nullptr;   // boundary staging, 6*nx*ny doubles
    double* d_full_b = nullptr;
};

template <typename T, int NX, int NY> struct Solver2DT : Solver2D {
    using C = typename ctype<T>::type;
    using XG = XGeo<T, NX>;
    using YG = YGeo<T, NY>;
    static constexpr int NXH = NX / 2 + 1;
    static constexpr size_t H = (size_t)NXH * NY;
    static constexpr size_t N = (size_t)NX * NY;

    DevBuf<T> csq_m, csq_p, om2_m, om2_p, aal, c2d2g, l2d, kxh, kyv;
    DevBuf<T> gam_m, gam_p, wall, fbdy7, chi, newgam, skewgam;
    DevBuf<T> gam_c, csq_c, om2_c;   // third reference depth
    DevBuf<C> twx, twy;
    DevBuf<T> src_sig, src_skew;
    DevBuf<double> sig_y, sig_c, skew_y, skew_c;
    std::vector<double> tsig;
    int nline = 0, nt = 0;
    DevBuf<T> eta_p, u_p, v_p;
    DevBuf<T> m0u_p, m0v_p;   // HS3 only
    DevBuf<C> work;     // 16 W slots + dH1 + dH2p
    DevBuf<double2> y, ynew;
    DevBuf<C> kbuf;     // 6 stage derivatives
    DevBuf<C> ytmp;
    DevBuf<unsigned long long> rmax;
    unsigned long long* h_rmax = nullptr;
    Rhs2DArgs<T> args{};
    // ---- breaking state
    int brk_on = 0;
    double delb = 1.5;
    double dxg = 1, dyg = 1;
    CrestMachine<T> cm;                    // criterion, crest lists, B (crest2d.h)
    DevBuf<T> flux, sbx, sby, s_eta;
    DevBuf<C> hilx;                        // Hilbert transform of eta: x-intermediate
    // The detection kernels (peaks, Hilbert transform, crest test) read only
    // the physical eta, u, v of the first x-pass and their results are waited
    // for at the end of the evaluation, so they run on a stream of their own
    // beside the rest of the pipeline (they fill the partial last round of
    // blocks of the y-passes).
    hipStream_t dstream = nullptr;
    hipEvent_t ev_fork = nullptr;
    DevBuf<T> hilp;                        // ... and physical
    bool sb_zero = true;                   // Sb^ known to be all zero
    bool last_gated = false;               // last RHS call skipped breaking (tprev >= t)
    DevBuf<C> sbh;                         // Sb_x^, Sb_y^ half spectra (2H)
    DevBuf<C> col0;                        // pre-mask eta derivative, column kx = 0 (NY)
    bool flat = false;
    bool hs1 = false;   // HS1 dynamics (linear AB1 right-hand side)
    bool hs3 = false;   // HS3 dynamics (AB3: the AB2 terms plus the third-order ones)
    bool three = false; // three reference depths (HSSgen_3depth)
    int src_on = 0;
    double g = 9.80665, eps_abs = 1e-3, eps_rel = 1e-2;
    // HIP events bracketing every RHS evaluation on the context's stream
    static constexpr int NEV = 64;
    hipEvent_t ev_a[NEV] = {}, ev_b[NEV] = {};
    int nev = 0;
    // ---- device step control (graph mode)
    DevBuf<DevStep> dstep;
    DevStep* h_dstep = nullptr;            // pinned copy of the device state
    DevBuf<double> tsig_d;                 // source sample times on the device
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    double2 *gA = nullptr, *gB = nullptr;  // the two state buffers the graph was captured with
    bool capturing = false;
    double cap_frac = 0.0;                 // stage time fraction of the RHS being captured
    hipEvent_t gev_a = nullptr, gev_b = nullptr;
    long long g_att = 0;                   // attempts counted into the stats so far
    std::vector<int> mcur;                 // members (nb > 1): buffer holding each one's state

    ~Solver2DT() override {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (h_dstep) (void)hipHostFree(h_dstep);
        if (gev_a) (void)hipEventDestroy(gev_a);
        if (gev_b) (void)hipEventDestroy(gev_b);
        if (h_rmax) (void)hipHostFree(h_rmax);
        for (int i = 0; i < NEV; ++i) {
            if (ev_a[i]) (void)hipEventDestroy(ev_a[i]);
            if (ev_b[i]) (void)hipEventDestroy(ev_b[i]);
        }
        if (d_full_a) (void)hipFree(d_full_a);
        if (d_full_b) (void)hipFree(d_full_b);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (dstream) (void)hipStreamDestroy(dstream);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // full-grid [ny][nx] -> half layout [kx][ky], kx <= nx/2
    static std::vector<T> to_half(const double* a, double scale = 1.0) {
        std::vector<T> out(H);
        for (int kx = 0; kx < NXH; ++kx)
            for (int ky = 0; ky < NY; ++ky) out[(size_t)kx * NY + ky] = (T)(a[(size_t)ky * NX + kx] * scale);
        return out;
    }
    static std::vector<T> to_phys(const double* a, double scale = 1.0) {
        std::vector<T> out(N);
        for (size_t i = 0; i < N; ++i) out[i] = (T)(a[i] * scale);
        return out;
    }
    int init(const hw_config_2d* c) override {
        nx = NX; ny = NY; prec = c->precision;
        flat = c->flat != 0;
        if (c->model < 0 || c->model > 3) {
            set_error("hw_config_2d.model must be 0 or 2 (HS2), 1 (HS1) or 3 (HS3), got %d", c->model);
            return HW_ERR_ARG;
        }
        hs1 = c->model == 1;
        hs3 = c->model == 3;
        g = c->grav; eps_abs = c->eps_abs; eps_rel = c->eps_rel;
        HW_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        for (int i = 0; i < NEV; ++i) {
            HW_TRY(hipEventCreate(&ev_a[i]));
            HW_TRY(hipEventCreate(&ev_b[i]));
        }
        int rc = 0;
        if (!c->kx || !c->ky || !c->aal || !c->fbdy || !c->wallchar || !c->chi) {
            set_error("hw_config_2d: kx, ky, aal, fbdy, wallchar and chi are required");
            return HW_ERR_ARG;
        }
        std::vector<T> kx(NXH), ky(NY);
        for (int i = 0; i < NXH; ++i) kx[i] = (T)c->kx[i];
        for (int i = 0; i < NY; ++i) ky[i] = (T)c->ky[i];
        if ((rc = kxh.upload(kx, stream)) || (rc = kyv.upload(ky, stream))) return rc;
        if ((rc = aal.upload(to_half(c->aal), stream))) return rc;
        if ((rc = fbdy7.upload(to_phys(c->fbdy, 7.0 * c->cpeak), stream))) return rc;
        if ((rc = wall.upload(to_phys(c->wallchar), stream))) return rc;
        if ((rc = chi.upload(to_phys(c->chi), stream))) return rc;
        if (flat) {
            if (!c->c2d || !c->l2d) { set_error("flat model needs c2d and l2d"); return HW_ERR_ARG; }
            std::vector<double> c2(N);
            for (size_t i = 0; i < N; ++i) c2[i] = c->c2d[i] * c->c2d[i] / c->grav;
            if ((rc = c2d2g.upload(to_half(c2.data()), stream)) || (rc = l2d.upload(to_half(c->l2d), stream))) return rc;
        } else {
            if (!c->gam_min || !c->gam_plus || !c->csq_min || !c->csq_plus || !c->om2_min || !c->om2_plus) {
                set_error("two-depth model needs gam/csq/om2 min and plus");
                return HW_ERR_ARG;
            }
            if ((rc = gam_m.upload(to_phys(c->gam_min), stream)) || (rc = gam_p.upload(to_phys(c->gam_plus), stream)) ||
                (rc = csq_m.upload(to_half(c->csq_min), stream)) || (rc = csq_p.upload(to_half(c->csq_plus), stream)) ||
                (rc = om2_m.upload(to_half(c->om2_min), stream)) || (rc = om2_p.upload(to_half(c->om2_plus), stream)))
                return rc;
            three = c->gam_mid != nullptr;
            if (three) {
                if (!c->csq_mid || !c->om2_mid) {
                    set_error("three reference depths need gam_mid, csq_mid and om2_mid");
                    return HW_ERR_ARG;
                }
                if ((rc = gam_c.upload(to_phys(c->gam_mid), stream)) || (rc = csq_c.upload(to_half(c->csq_mid), stream)) ||
                    (rc = om2_c.upload(to_half(c->om2_mid), stream)))
                    return rc;
            }
        }
        if ((rc = twx.upload(x_twiddles<T, NX>(), stream)) || (rc = twy.upload(y_twiddles<T, NY>(), stream))) return rc;
        src_on = c->source_on;
        if (src_on) {
            nline = c->nline; nt = c->nt;
            int want = c->vertical ? NY : NX;
            if (nline != want || nt < 3 || !c->t_sig || !c->sig || !c->skew_sig || !c->new_gam || !c->skew_gam) {
                set_error("source: nline must be %d and t_sig/sig/skew_sig/new_gam/skew_gam given (nt>=3)", want);
                return HW_ERR_ARG;
            }
            tsig.assign(c->t_sig, c->t_sig + nt);
            // natural-spline half second derivatives per line (gsl cspline), tables [nt][nline]
            std::vector<double> ty((size_t)nt * nline), tc((size_t)nt * nline), ky_((size_t)nt * nline), kc((size_t)nt * nline);
            std::vector<double> yy(nt), cc(nt);
            for (int s = 0; s < 2; ++s) {
                const double* src = s == 0 ? c->sig : c->skew_sig;
                for (int j = 0; j < nline; ++j) {
                    for (int n = 0; n < nt; ++n) yy[n] = src[(size_t)j * nt + n];
                    spline_half_second(tsig.data(), yy.data(), nt, cc.data());
                    for (int n = 0; n < nt; ++n) {
                        (s == 0 ? ty : ky_)[(size_t)n * nline + j] = yy[n];
                        (s == 0 ? tc : kc)[(size_t)n * nline + j] = cc[n];
                    }
                }
            }
            if ((rc = sig_y.upload(ty, stream)) || (rc = sig_c.upload(tc, stream)) || (rc = skew_y.upload(ky_, stream)) ||
                (rc = skew_c.upload(kc, stream)))
                return rc;
            if ((rc = src_sig.alloc(nline)) || (rc = src_skew.alloc(nline))) return rc;
            if ((rc = newgam.upload(to_phys(c->new_gam), stream)) || (rc = skewgam.upload(to_phys(c->skew_gam), stream))) return rc;
        }
        tprev = c->t0;
        brk_on = c->breaking_on && !hs1;   // rhs_AB1_* never calls breaking_process
        if (nb < 1) { set_error("members must be >= 1"); return HW_ERR_ARG; }
        if ((rc = s_eta.alloc((size_t)NY * nb))) return rc;
        if (brk_on) {
            if ((rc = cm.init(c, NX, NY, stream))) return rc;
            HW_TRY(hipStreamCreateWithFlags(&dstream, hipStreamNonBlocking));
            HW_TRY(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
            delb = c->delb;
            dxg = cm.dxg; dyg = cm.dyg;
            if ((rc = flux.alloc(N)) || (rc = sbx.alloc(N)) || (rc = sby.alloc(N)) || (rc = hilx.alloc(H)) ||
                (rc = hilp.alloc(N)) || (rc = sbh.alloc(2 * H)) || (rc = col0.alloc(NY)))
                return rc;
            k_fill<C><<<grid1(2 * H), 256, 0, stream>>>(sbh.p, cmk_h((T)0, (T)0), 2 * H);
            HW_LAUNCH_CHECK();
        }
        if (nb > 1 && (brk_on || src_on || three)) {
            set_error("a context with several members needs breaking, the influx source and a third reference "
                      "depth off");
            return HW_ERR_ARG;
        }
        const size_t B = (size_t)nb;
        if ((rc = eta_p.alloc(N * B)) || (rc = u_p.alloc(N * B)) || (rc = v_p.alloc(N * B))) return rc;
        if ((hs3 || three) && ((rc = m0u_p.alloc(N * B)) || (rc = m0v_p.alloc(N * B)))) return rc;
        if ((rc = work.alloc(18 * H * B))) return rc;
        if ((rc = y.alloc(3 * H * B)) || (rc = ynew.alloc(3 * H * B)) || (rc = kbuf.alloc(6 * 3 * H * B)) ||
            (rc = ytmp.alloc(3 * H * B)))
            return rc;
        mcur.assign(B, 0);
        if ((rc = rmax.alloc(1))) return rc;
        HW_TRY(hipHostMalloc((void**)&h_rmax, sizeof(unsigned long long), hipHostMallocDefault));
        HW_TRY(hipMemsetAsync(y.p, 0, sizeof(double2) * 3 * H * B, stream));
        // fixed kernel arguments
        args.csq_m = csq_m.p; args.csq_p = csq_p.p; args.om2_m = om2_m.p; args.om2_p = om2_p.p;
        args.aal = aal.p; args.c2d2g = c2d2g.p; args.l2d = l2d.p; args.kxh = kxh.p; args.ky = kyv.p;
        args.gam_m = gam_m.p; args.gam_p = gam_p.p; args.wallchar = wall.p; args.fbdy7 = fbdy7.p; args.chi = chi.p;
        args.new_gam = newgam.p; args.skew_gam = skewgam.p;
        args.src_sig = src_sig.p; args.src_skew = src_skew.p;
        args.src_on = src_on; args.vertical = c->vertical; args.propagation = c->propagation;
        args.sbx = nullptr; args.sby = nullptr;
        args.eta_p = eta_p.p; args.u_p = u_p.p; args.v_p = v_p.p;
        args.m0u_p = m0u_p.p; args.m0v_p = m0v_p.p;
        args.gam_c = gam_c.p; args.csq_c = csq_c.p; args.om2_c = om2_c.p;
        for (int i = 0; i < 16; ++i) args.W[i] = work.p + (size_t)i * H;
        args.dH1 = work.p + 16 * H;
        args.dH2p = work.p + 17 * H;
        args.twx = twx.p; args.twy = twy.p;
        args.g = (T)g;
        args.s_eta_out = s_eta.p;
        args.zs_state = 3 * H; args.zs_work = 18 * H; args.zs_phys = N; args.zs_row = NY;
        HW_TRY(hipStreamSynchronize(stream));
        return HW_OK;
    }

    static C cmk_h(T a, T b) { C r; r.x = a; r.y = b; return r; }

    // accumulate the recorded RHS durations into st.rhs_ms_total
    int collect_timing() {
        for (int i = 0; i < nev; ++i) {
            HW_TRY(hipEventSynchronize(ev_b[i]));
            float ms = 0.f;
            HW_TRY(hipEventElapsedTime(&ms, ev_a[i], ev_b[i]));
            st.rhs_ms_total += ms;
        }
        nev = 0;
        return HW_OK;
    }

    // ------------------------------------------------------------ RHS --
    // every HW_RHS_TIME_EVERY-th evaluation is bracketed by HIP events
    static constexpr long HW_RHS_TIME_EVERY = 4;
    int rhs(double t, const C* in, C* out) {
        const bool timed = !capturing && st.rhs_evals % HW_RHS_TIME_EVERY == 0;
        if (timed) {
            if (nev == NEV) {
                int rc = collect_timing();
                if (rc) return rc;
            }
            HW_TRY(hipEventRecord(ev_a[nev], stream));
        }
        args.eta_hat = in; args.u_hat = in + H; args.v_hat = in + 2 * H;
        args.k_eta = out; args.k_u = out + H; args.k_v = out + 2 * H;
        args.sp_n = 0;
        if (src_on && capturing) {
            k_src_eval_g<T><<<(nline + 255) / 256, 256, 0, stream>>>(dstep.p, cap_frac, tsig_d.p, nt, sig_y.p, sig_c.p,
                                                                     skew_y.p, skew_c.p, nline, src_sig.p, src_skew.p);
            HW_LAUNCH_CHECK();
        } else if (src_on) {
            // the line signals at t: evaluated by the first y-pass
            int idx = bsearch_interval(tsig, t);
            args.sp_y = sig_y.p; args.sp_c = sig_c.p; args.sk_y = skew_y.p; args.sk_c = skew_c.p;
            args.sp_n = nline;
            args.sp_idx = idx;
            args.sp_dx = t - tsig[idx];
            args.sp_hh = tsig[idx + 1] - tsig[idx];
            args.src_sig_w = src_sig.p; args.src_skew_w = src_skew.p;
        }
        const dim3 gx(NY / XG::RR, 1, nb), bx(XG::NT);
        const dim3 gy((NXH + YG::CC - 1) / YG::CC, 1, nb), by(YG::NT);
        // 2d/fun_odesolver.c:4580-4587: breaking runs only when tprev < t
        const bool brk_now = brk_on && tprev < t;
        args.eta_col0 = brk_now ? col0.p : nullptr;
        // the peak-search counters are zeroed by the first y-pass
        args.zero_cnt = brk_now ? cm.pkcount : nullptr;
        // the middle reference depth's share of each HSS (HSSgen_3depth)
        auto mid_m0 = [&]() {
            k3d_y1<T, NX, NY><<<gy, by, 0, stream>>>(args);
            k3d_x2<T, NX, NY><<<gx, bx, 0, stream>>>(args);
            k3d_y3<T, NX, NY><<<gy, by, 0, stream>>>(args);
        };
        auto mid_lem0_a = [&]() {
            k3d_x4<T, NX, NY><<<gx, bx, 0, stream>>>(args);
            k3d_y5<T, NX, NY><<<gy, by, 0, stream>>>(args);
        };
        auto mid_lem0_b = [&]() {
            k3d_x6<T, NX, NY><<<gx, bx, 0, stream>>>(args);
            k_add_x<C><<<grid1(H), 256, 0, stream>>>(args.W[S_ACC1], args.W[E3_ACC], H);
        };
        if (hs1) {
            // rhs_AB1_2DB_uv / rhs_AB1_2DF_uv (2d/fun_odesolver.c:4350-4456)
            if (!flat) {
                k2b_y1<T, NX, NY><<<gy, by, 0, stream>>>(args);
                k2b_x1<T, NX, NY><<<gx, bx, 0, stream>>>(args);
                k2b_y2<T, NX, NY><<<gy, by, 0, stream>>>(args);
                if (three) mid_m0();
            } else {
                k2f_y1<T, NX, NY><<<gy, by, 0, stream>>>(args);
            }
            k1_x<T, NX, NY><<<gx, bx, 0, stream>>>(args, flat ? 0 : 1);
            k1_y<T, NX, NY><<<gy, by, 0, stream>>>(args, flat ? 1 : 0);
        } else if (hs3 && !flat) {
            // rhs_AB3_2DB_uv (2d/fun_odesolver.c:4735-4843)
            k2b_y1<T, NX, NY><<<gy, by, 0, stream>>>(args);
            k2b_x1<T, NX, NY><<<gx, bx, 0, stream>>>(args);
            if (brk_now) { int rc = breaking_detect(); if (rc) return rc; }
            k2b_y2<T, NX, NY><<<gy, by, 0, stream>>>(args);
            if (three) mid_m0();
            k2b_x3<T, NX, NY, true><<<gx, bx, 0, stream>>>(args);
            k2b_y4<T, NX, NY><<<gy, by, 0, stream>>>(args);
            if (three) mid_lem0_a();
            k3b_x5a<T, NX, NY><<<gx, bx, 0, stream>>>(args);
            if (three) mid_lem0_b();
            k3_ya<T, NX, NY, false><<<gy, by, 0, stream>>>(args);
            if (three) {
                k3_xb<T, NX, NY, false, true><<<gx, bx, 0, stream>>>(args);
                k3_yc<T, NX, NY, false, true><<<gy, by, 0, stream>>>(args);
                k3b_x5b<T, NX, NY, true><<<gx, bx, 0, stream>>>(args);
            } else {
                k3_xb<T, NX, NY, false><<<gx, bx, 0, stream>>>(args);
                k3_yc<T, NX, NY, false><<<gy, by, 0, stream>>>(args);
                k3b_x5b<T, NX, NY><<<gx, bx, 0, stream>>>(args);
            }
            k2b_y6<T, NX, NY, true><<<gy, by, 0, stream>>>(args);
            k2b_x7<T, NX, NY><<<gx, bx, 0, stream>>>(args);
            k2b_y8<T, NX, NY><<<gy, by, 0, stream>>>(args);
        } else if (hs3) {
            // rhs_AB3_2DF_uv (2d/fun_odesolver.c:4628-4733)
            k2f_y1<T, NX, NY><<<gy, by, 0, stream>>>(args);
            k2f_x2<T, NX, NY, true><<<gx, bx, 0, stream>>>(args);
            if (brk_now) { int rc = breaking_detect(); if (rc) return rc; }
            k2f_y3<T, NX, NY><<<gy, by, 0, stream>>>(args);
            k3_ya<T, NX, NY, true><<<gy, by, 0, stream>>>(args);
            k3_xb<T, NX, NY, true><<<gx, bx, 0, stream>>>(args);
            k3_yc<T, NX, NY, true><<<gy, by, 0, stream>>>(args);
            k2f_x4<T, NX, NY><<<gx, bx, 0, stream>>>(args);
            k2f_y5<T, NX, NY><<<gy, by, 0, stream>>>(args);
        } else if (!flat) {
            k2b_y1<T, NX, NY><<<gy, by, 0, stream>>>(args);
            k2b_x1<T, NX, NY><<<gx, bx, 0, stream>>>(args);
            if (brk_now) { int rc = breaking_detect(); if (rc) return rc; }
            k2b_y2<T, NX, NY><<<gy, by, 0, stream>>>(args);
            if (three) {
                mid_m0();
                k2b_x3<T, NX, NY, true><<<gx, bx, 0, stream>>>(args);
                k2b_y4<T, NX, NY><<<gy, by, 0, stream>>>(args);
                mid_lem0_a();
                k2b_x5<T, NX, NY><<<gx, bx, 0, stream>>>(args);
                mid_lem0_b();
            } else {
                k2b_x3<T, NX, NY><<<gx, bx, 0, stream>>>(args);
                k2b_y4<T, NX, NY><<<gy, by, 0, stream>>>(args);
                k2b_x5<T, NX, NY><<<gx, bx, 0, stream>>>(args);
            }
            k2b_y6<T, NX, NY><<<gy, by, 0, stream>>>(args);
            k2b_x7<T, NX, NY><<<gx, bx, 0, stream>>>(args);
            k2b_y8<T, NX, NY><<<gy, by, 0, stream>>>(args);
        } else {
            k2f_y1<T, NX, NY><<<gy, by, 0, stream>>>(args);
            k2f_x2<T, NX, NY><<<gx, bx, 0, stream>>>(args);
            if (brk_now) { int rc = breaking_detect(); if (rc) return rc; }
            k2f_y3<T, NX, NY><<<gy, by, 0, stream>>>(args);
            k2f_x4<T, NX, NY><<<gx, bx, 0, stream>>>(args);
            k2f_y5<T, NX, NY><<<gy, by, 0, stream>>>(args);
        }
        HW_LAUNCH_CHECK();
        if (brk_on) {
            // the (possibly stale) Sb^ is added in every call
            if (brk_now) {
                // (fun_sbreaking's in-place projection of delphiH, which the
                // eta derivative keeps, is applied by the last y-pass)
                int rc = breaking(t);
                if (rc) return rc;
            }
            last_gated = !brk_now;
            if (!sb_zero) {
                k_add_sb<T><<<grid1(H), 256, 0, stream>>>(args.k_u, args.k_v, sbh.p, sbh.p + H, aal.p, H);
                HW_LAUNCH_CHECK();
            }
        }
        if (timed) {
            HW_TRY(hipEventRecord(ev_b[nev], stream));
            nev++;
            st.rhs_timed++;
        }
        if (!capturing) st.rhs_evals++;
        return HW_OK;
    }

    // ------------------------------------------------------- breaking --
    // breaking_process (2d/breaking.c:505-733) + fun_sbreaking
    // (2d/fun_odesolver.c:985-1019) for one RHS call at time t.
    // Peaks, Hilbert transform and crest test (breaking_criterion) need only
    // eta, u, v of this call, so they are launched right after the first
    // x-pass; the counts and the first crest candidates come back with an
    // asynchronous copy while the rest of the pipeline runs.
    int breaking_detect() {
        const dim3 gx(NY / XG::RR), bx(XG::NT);
        const dim3 gy((NXH + YG::CC - 1) / YG::CC), by(YG::NT);
        int rc;
        // fork after the first x-pass; CrestMachine::run joins (host wait on
        // its event, then the main stream waits on it too)
        HW_TRY(hipEventRecord(ev_fork, stream));
        HW_TRY(hipStreamWaitEvent(dstream, ev_fork, 0));
        if ((rc = cm.detect_begin(dstream, eta_p.p, true))) return rc;
        k_brk_hilb_y<T, NX, NY><<<gy, by, 0, dstream>>>(args, hilx.p, cm.pkcount);
        k_t_xinv<T, NX, NY><<<gx, bx, 0, dstream>>>(hilx.p, hilp.p, nullptr, twx.p, cm.pkcount);
        HW_LAUNCH_CHECK();
        return cm.detect_end(dstream, eta_p.p, hilp.p, u_p.p, v_p.p);
    }

    int breaking(double t) {
        const dim3 gx(NY / XG::RR), bx(XG::NT);
        const dim3 gy((NXH + YG::CC - 1) / YG::CC), by(YG::NT);
        bool have_b = false;
        int rc = cm.run(t, stream, eta_p.p, u_p.p, v_p.p, &have_b);
        if (rc) return rc;
        if (!have_b) {
            // B == 0 everywhere: zero flux, zero source
            if (!sb_zero) k_fill<C><<<grid1(2 * H), 256, 0, stream>>>(sbh.p, cmk_h((T)0, (T)0), 2 * H);
            HW_LAUNCH_CHECK();
            sb_zero = true;
            return HW_OK;
        }
        sb_zero = false;
        // ---- fun_sbreaking: delphi, flux, clipped gradient, transforms
        C* Xg = args.W[12];
        C* Xd = args.W[13];
        k_brk_g_x<T, NX, NY><<<gx, bx, 0, stream>>>(args, s_eta.p, Xg);
        k_brk_del_y<T, NX, NY><<<gy, by, 0, stream>>>(args, flat ? 1 : 0, Xg, Xd);
        k_brk_flux_x<T, NX, NY><<<gx, bx, 0, stream>>>(args, args.W[flat ? (int)F_AETA : (int)S_AETA], Xd, cm.Bf.p, cm.depth.p,
                                                      (T)(delb * delb), flux.p);
        k_brk_sb<T><<<dim3((NX + 255) / 256, NY), 256, 0, stream>>>(flux.p, eta_p.p, cm.depth.p, dxg, dyg, NX, NY, sbx.p, sby.p);
        k_brk_sb_x<T, NX, NY><<<gx, bx, 0, stream>>>(args, sbx.p, sby.p, args.W[10], args.W[11]);
        k_brk_sb_y<T, NX, NY><<<gy, by, 0, stream>>>(args, args.W[10], args.W[11], sbh.p, sbh.p + H);
        HW_LAUNCH_CHECK();
        return HW_OK;
    }

    int breaking_state(double* Bout, int* fl) override {
        if (fl) *fl = cm.flagbr;
        if (Bout) {
            // the reference clears B on every call and fills it only when
            // breaking runs (2d/fun_odesolver.c:4581-4585)
            if (!brk_on || last_gated) {
                for (size_t i = 0; i < N; ++i) Bout[i] = 0.0;
            } else {
                std::vector<T> hb(N);
                HW_TRY(hipMemcpyAsync(hb.data(), cm.Bf.p, sizeof(T) * N, hipMemcpyDeviceToHost, stream));
                HW_TRY(hipStreamSynchronize(stream));
                for (size_t i = 0; i < N; ++i) Bout[i] = hb[i];
            }
        }
        return HW_OK;
    }

    static dim3 grid1(size_t n) {
        size_t b = (n + 255) / 256;
        if (b > 65535u * 8) b = 65535u * 8;
        return dim3((unsigned)b);
    }

    int rhs_full_dev(double t, const double* d_in, double* d_out) override {
        dim3 b(32, 8), gi((NXH + 31) / 32, (NY + 31) / 32, 3), go((NX + 31) / 32, (NY + 31) / 32, 3);
        k_full_to_half<C><<<gi, b, 0, stream>>>(d_in, ytmp.p, NX, NY);
        HW_LAUNCH_CHECK();
        int rc = rhs(0.0 + t, ytmp.p, kbuf.p);
        if (rc) return rc;
        k_half_to_full<C><<<go, b, 0, stream>>>(kbuf.p, d_out, NX, NY);
        HW_LAUNCH_CHECK();
        HW_TRY(hipStreamSynchronize(stream));
        return collect_timing();
    }

    int set_state_dev(const double* d_full) override {
        dim3 b(32, 8), gi((NXH + 31) / 32, (NY + 31) / 32, 3);
        k_full_to_half<double2><<<gi, b, 0, stream>>>(d_full, y.p, NX, NY);
        HW_LAUNCH_CHECK();
        HW_TRY(hipStreamSynchronize(stream));
        return HW_OK;
    }

    int get_state_dev(double* d_full) override {
        dim3 b(32, 8), go((NX + 31) / 32, (NY + 31) / 32, 3);
        k_half_to_full<double2><<<go, b, 0, stream>>>(y.p, d_full, NX, NY);
        HW_LAUNCH_CHECK();
        HW_TRY(hipStreamSynchronize(stream));
        return HW_OK;
    }

    int stage(double h, std::initializer_list<int> ks, std::initializer_list<double> bs) {
        StageCoef sc{};
        sc.nk = 0;
        auto bi = bs.begin();
        for (int kidx : ks) {
            sc.k[sc.nk] = K(kidx);
            sc.b[sc.nk] = *bi++;
            sc.nk++;
        }
        k_rk_stage<T><<<grid1(3 * H), 256, 0, stream>>>(y.p, sc, h, ytmp.p, 3 * H);
        HW_LAUNCH_CHECK();
        return HW_OK;
    }

    C* K(int i) { return kbuf.p + (size_t)i * 3 * H * nb; }

    // One attempted rkf45 step from (t0, y) with step h0, enqueued on the
    // stream: K0 = f(t0, y) first when need_k0 (gsl's dydt_in, computed once
    // per evolve_apply call and reused by retried attempts), the five further
    // stages, the error norm and, with breaking, rkf45's dydt_out evaluation.
    // The norm's copy to the host is asynchronous; attempt_result waits for it.
    int attempt_issue(double t0, double h0, bool need_k0) override {
        int rc;
        if (need_k0) {
            if ((rc = stage(0.0, {}, {})) || (rc = rhs(t0, ytmp.p, K(0)))) return rc;
        }
        if ((rc = stage(h0, {0}, {RK_AH[0]})) || (rc = rhs(t0 + RK_AH[0] * h0, ytmp.p, K(1)))) return rc;
        if ((rc = stage(h0, {0, 1}, {RK_B3[0], RK_B3[1]})) || (rc = rhs(t0 + RK_AH[1] * h0, ytmp.p, K(2)))) return rc;
        if ((rc = stage(h0, {0, 1, 2}, {RK_B4[0], RK_B4[1], RK_B4[2]})) || (rc = rhs(t0 + RK_AH[2] * h0, ytmp.p, K(3)))) return rc;
        if ((rc = stage(h0, {0, 1, 2, 3}, {RK_B5[0], RK_B5[1], RK_B5[2], RK_B5[3]})) ||
            (rc = rhs(t0 + RK_AH[3] * h0, ytmp.p, K(4))))
            return rc;
        if ((rc = stage(h0, {0, 1, 2, 3, 4}, {RK_B6[0], RK_B6[1], RK_B6[2], RK_B6[3], RK_B6[4]})) ||
            (rc = rhs(t0 + RK_AH[4] * h0, ytmp.p, K(5))))
            return rc;
        k_zero_u64<<<1, 64, 0, stream>>>(rmax.p, 1);
        // with breaking, the final kernel also writes y_new in T for the
        // dydt_out evaluation below
        k_rk_final<T><<<dim3(2048), 256, 0, stream>>>(y.p, K(0), K(2), K(3), K(4), K(5), h0, eps_abs, eps_rel, ynew.p,
                                                        rmax.p, 3 * H, brk_on ? ytmp.p : nullptr);
        HW_LAUNCH_CHECK();
        if (brk_on) {
            // rkf45's dydt_out = f(t + h, y_new): its breaking side effects
            // are part of the reference's sequence (K(1) is free here)
            if ((rc = rhs(t0 + h0, ytmp.p, K(1)))) return rc;
        }
        HW_TRY(hipMemcpyAsync(h_rmax, rmax.p, sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
        return HW_OK;
    }

    // max_i E_i / D_i of the last issued attempt (GSL starts rmax at DBL_MIN)
    int attempt_result(double* rm) override {
        HW_TRY(hipStreamSynchronize(stream));
        std::memcpy(rm, h_rmax, sizeof(double));
        if (*rm < 2.2250738585072014e-308) *rm = 2.2250738585072014e-308;
        return HW_OK;
    }

    // the last attempt is accepted: y <- y_new at time t
    int accept_step(double t) override {
        std::swap(y.p, ynew.p);
        st.steps_accepted++;
        tprev = t;   // ode_solver_uv: tprev = t after each evolve_apply
        return collect_timing();
    }

    // gsl_odeiv2_evolve_apply with rkf45 + control_y, restated (see header).
    // The rkf45 stepper's dydt_out evaluation is not performed without
    // breaking: control_y does not read it and the RHS has no other side
    // effects on this path.
    int evolve_apply(double* t, double t1, double* h) override {
        const double t0 = *t;
        double h0 = *h;
        const double dt = t1 - t0;
        if ((dt < 0.0 && h0 > 0.0) || (dt > 0.0 && h0 < 0.0)) {
            set_error("step direction must match interval direction");
            return HW_ERR_ARG;
        }
        bool need_k0 = true;
        for (;;) {
            bool final_step = false;
            if ((dt >= 0.0 && h0 > dt) || (dt < 0.0 && h0 < dt)) { h0 = dt; final_step = true; }
            int rc;
            double rm;
            if ((rc = attempt_issue(t0, h0, need_k0)) || (rc = attempt_result(&rm))) return rc;
            need_k0 = false;
            *t = final_step ? t1 : t0 + h0;
            double h_old = h0;
            int adj = hadjust_rule(rm, &h0);
            if (adj < 0) {
                double t_curr = *t, t_next = *t + h0;
                if (std::fabs(h0) < std::fabs(h_old) && t_next != t_curr) {
                    st.steps_rejected++;
                    continue;
                }
                *h = h0;
                *t = t0;
                set_error("step size underflow at t=%g", t0);
                return HW_ERR_STEP;
            }
            *h = h0;
            return accept_step(*t);
        }
    }

    // ------------------------------------------- device step control --
    // One attempt of evolve_apply as graph nodes: the stage times come from
    // the device state (k_step_begin), K0 is evaluated in every attempt (a
    // retried attempt recomputes the same K0 the host loop keeps), and the
    // error norm goes to k_step_end, which accepts or rejects on the device.
    // With several members every launch covers all of them (grid.z or the
    // element range), each member with its own t, h and accept decision.
    int stage_g(std::initializer_list<int> ks, std::initializer_list<double> bs) {
        StageCoef sc{};
        sc.nk = 0;
        auto bi = bs.begin();
        for (int kidx : ks) {
            sc.k[sc.nk] = K(kidx);
            sc.b[sc.nk] = *bi++;
            sc.nk++;
        }
        k_rk_stage_g<T><<<grid1(3 * H * nb), 256, 0, stream>>>(gA, gB, dstep.p, sc, ytmp.p, 3 * H, (size_t)nb * 3 * H);
        HW_LAUNCH_CHECK();
        return HW_OK;
    }

    int attempt_body_g() {
        int rc;
        const unsigned sb = (unsigned)((nb + 63) / 64);
        k_step_begin<<<sb, 64, 0, stream>>>(dstep.p, nb);
        HW_LAUNCH_CHECK();
        cap_frac = 0.0;
        if ((rc = stage_g({}, {})) || (rc = rhs(0.0, ytmp.p, K(0)))) return rc;
        cap_frac = RK_AH[0];
        if ((rc = stage_g({0}, {RK_AH[0]})) || (rc = rhs(0.0, ytmp.p, K(1)))) return rc;
        cap_frac = RK_AH[1];
        if ((rc = stage_g({0, 1}, {RK_B3[0], RK_B3[1]})) || (rc = rhs(0.0, ytmp.p, K(2)))) return rc;
        cap_frac = RK_AH[2];
        if ((rc = stage_g({0, 1, 2}, {RK_B4[0], RK_B4[1], RK_B4[2]})) || (rc = rhs(0.0, ytmp.p, K(3)))) return rc;
        cap_frac = RK_AH[3];
        if ((rc = stage_g({0, 1, 2, 3}, {RK_B5[0], RK_B5[1], RK_B5[2], RK_B5[3]})) || (rc = rhs(0.0, ytmp.p, K(4))))
            return rc;
        cap_frac = RK_AH[4];
        if ((rc = stage_g({0, 1, 2, 3, 4}, {RK_B6[0], RK_B6[1], RK_B6[2], RK_B6[3], RK_B6[4]})) ||
            (rc = rhs(0.0, ytmp.p, K(5))))
            return rc;
        // one block row per member, so that each block's norm belongs to one member
        const unsigned fb = nb == 1 ? 2048u : (unsigned)std::max<size_t>(1, std::min<size_t>(2048, (3 * H + 255) / 256));
        k_rk_final_g<T><<<dim3(fb, nb), 256, 0, stream>>>(gA, gB, dstep.p, K(0), K(2), K(3), K(4), K(5), 3 * H);
        HW_LAUNCH_CHECK();
        k_step_end<<<sb, 64, 0, stream>>>(dstep.p, nb);
        HW_LAUNCH_CHECK();
        return HW_OK;
    }

    int capture_attempt() {
        int rc;
        if (!dstep.p && (rc = dstep.alloc((size_t)nb))) return rc;
        if (!h_dstep) HW_TRY(hipHostMalloc((void**)&h_dstep, sizeof(DevStep) * nb, hipHostMallocDefault));
        if (src_on && !tsig_d.p && (rc = tsig_d.upload(tsig, stream))) return rc;
        if (!gev_a) {
            HW_TRY(hipEventCreate(&gev_a));
            HW_TRY(hipEventCreate(&gev_b));
        }
        if ((rc = collect_timing())) return rc;
        HW_TRY(hipStreamSynchronize(stream));
        gA = y.p;
        gB = ynew.p;
        HW_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        capturing = true;
        rc = attempt_body_g();
        capturing = false;
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamEndCapture(stream, &gr);
        if (rc) {
            if (gr) (void)hipGraphDestroy(gr);
            return rc;
        }
        HW_TRY(e);
        graph = gr;
        HW_TRY(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        return HW_OK;
    }

    int graph_begin(const double* t, double t1, const double* h) override {
        if (brk_on) {
            set_error("device step control (HW_STEP_GRAPH) needs breaking off");
            return HW_ERR_ARG;
        }
        int rc;
        if (!gexec && (rc = capture_attempt())) return rc;
        if (!((y.p == gA && ynew.p == gB) || (y.p == gB && ynew.p == gA))) {
            set_error("state buffers differ from the captured ones");
            return HW_ERR_HIP;
        }
        for (int z = 0; z < nb; ++z) {
            DevStep d{};
            d.t = t[z]; d.h = h[z]; d.t1 = t1;
            d.eps_abs = eps_abs; d.eps_rel = eps_rel;
            d.cur = nb == 1 ? (y.p == gB ? 1 : 0) : mcur[z];
            d.done = t[z] < t1 ? 0 : 1;
            h_dstep[z] = d;
        }
        g_att = 0;
        HW_TRY(hipMemcpyAsync(dstep.p, h_dstep, sizeof(DevStep) * nb, hipMemcpyHostToDevice, stream));
        return HW_OK;
    }

    int graph_launch() override {
        // about as many attempts as accepted steps remain at the current h
        // of the member with the most left
        double rem = 0.0;
        for (int z = 0; z < nb; ++z) {
            const DevStep& d = h_dstep[z];
            if (!d.done) rem = std::max(rem, (d.t1 - d.t) / d.h);
        }
        const int nl = rem >= 64.0 ? 64 : rem >= 1.0 ? (int)rem : 1;
        HW_TRY(hipEventRecord(gev_a, stream));
        for (int i = 0; i < nl; ++i) HW_TRY(hipGraphLaunch(gexec, stream));
        HW_TRY(hipEventRecord(gev_b, stream));
        HW_TRY(hipMemcpyAsync(h_dstep, dstep.p, sizeof(DevStep) * nb, hipMemcpyDeviceToHost, stream));
        return HW_OK;
    }

    // rhs_ms_total then covers whole attempts (stages, norm and control
    // too) of all members; rhs_evals counts member evaluations
    int graph_poll(bool* done) override {
        HW_TRY(hipStreamSynchronize(stream));
        long long att = 0;
        bool all = true;
        for (int z = 0; z < nb; ++z) {
            att += h_dstep[z].attempts;
            all = all && h_dstep[z].done != 0;
        }
        const long long natt = att - g_att;
        g_att = att;
        st.rhs_evals += 6 * natt;
        float ms = 0.f;
        HW_TRY(hipEventElapsedTime(&ms, gev_a, gev_b));
        st.rhs_ms_total += ms;
        st.rhs_timed += 6 * natt;
        *done = all;
        return HW_OK;
    }

    int graph_finish(double* t, double* h) override {
        int err = -1;
        for (int z = 0; z < nb; ++z) {
            const DevStep& d = h_dstep[z];
            st.steps_accepted += d.accepted;
            st.steps_rejected += d.rejected;
            t[z] = d.t;
            h[z] = d.h;
            if (d.err && err < 0) err = z;
            if (nb > 1) mcur[z] = d.cur;
        }
        if (nb == 1) {
            y.p = h_dstep[0].cur ? gB : gA;
            ynew.p = h_dstep[0].cur ? gA : gB;
            tprev = t[0];
        }
        if (err >= 0) {
            set_error("member %d: step size underflow at t=%g", err, h_dstep[err].t);
            return HW_ERR_STEP;
        }
        return HW_OK;
    }

    // members: the state of member z lives in the buffer mcur[z] names
    double2* member_ptr(int z) {
        double2* base = (mcur[z] ? ynew.p : y.p);
        return base + (size_t)z * 3 * H;
    }
    int set_member_dev(const double* d_full, int z) override {
        if (z < 0 || z >= nb) { set_error("member %d out of range", z); return HW_ERR_ARG; }
        dim3 b(32, 8), gi((NXH + 31) / 32, (NY + 31) / 32, 3);
        k_full_to_half<double2><<<gi, b, 0, stream>>>(d_full, member_ptr(z), NX, NY);
        HW_LAUNCH_CHECK();
        HW_TRY(hipStreamSynchronize(stream));
        return HW_OK;
    }
    int get_member_dev(double* d_full, int z) override {
        if (z < 0 || z >= nb) { set_error("member %d out of range", z); return HW_ERR_ARG; }
        dim3 b(32, 8), go((NX + 31) / 32, (NY + 31) / 32, 3);
        k_half_to_full<double2><<<go, b, 0, stream>>>(member_ptr(z), d_full, NX, NY);
        HW_LAUNCH_CHECK();
        HW_TRY(hipStreamSynchronize(stream));
        return HW_OK;
    }

    int fields_dev(double* eta, double* u, double* v) override {
        const dim3 gx(NY / XG::RR), bx(XG::NT);
        const dim3 gy((NXH + YG::CC - 1) / YG::CC), by(YG::NT);
        double* outs[3] = {eta, u, v};
        k_convert<double2, C><<<grid1(3 * H), 256, 0, stream>>>(y.p, ytmp.p, 3 * H);
        HW_LAUNCH_CHECK();
        for (int f = 0; f < 3; ++f) {
            if (!outs[f]) continue;
            k_t_yinv<T, NX, NY><<<gy, by, 0, stream>>>(ytmp.p + (size_t)f * H, args.W[0], twy.p);
            k_t_xinv<T, NX, NY><<<gx, bx, 0, stream>>>(args.W[0], eta_p.p, nullptr, twx.p);
            HW_LAUNCH_CHECK();
            k_convert_scalar<T, double><<<grid1(N), 256, 0, stream>>>(eta_p.p, outs[f], N);
            HW_LAUNCH_CHECK();
        }
        HW_TRY(hipStreamSynchronize(stream));
        return HW_OK;
    }
};

// stand-alone transforms (hw_fft2_forward / hw_fft2_inverse)
template <typename T, int NX, int NY>
int fft2_forward(const double* phys, double* re, double* im) {
    using C = typename ctype<T>::type;
    constexpr int NXH = NX / 2 + 1;
    constexpr size_t H = (size_t)NXH * NY, N = (size_t)NX * NY;
    std::vector<T> hp(N);
    for (size_t i = 0; i < N; ++i) hp[i] = (T)phys[i];
    DevBuf<T> dp;
    DevBuf<C> X, S, tx, ty;
    int rc;
    if ((rc = dp.upload(hp, 0)) || (rc = X.alloc(H)) || (rc = S.alloc(H)) ||
        (rc = tx.upload(x_twiddles<T, NX>(), 0)) || (rc = ty.upload(y_twiddles<T, NY>(), 0)))
        return rc;
    k_t_xfwd<T, NX, NY><<<NY / XGeo<T, NX>::RR, XGeo<T, NX>::NT>>>(dp.p, X.p, tx.p);
    k_t_yfwd<T, NX, NY><<<(NXH + YGeo<T, NY>::CC - 1) / YGeo<T, NY>::CC, YGeo<T, NY>::NT>>>(X.p, S.p, ty.p);
    HW_LAUNCH_CHECK();
    std::vector<C> hs(H);
    HW_TRY(hipMemcpy(hs.data(), S.p, sizeof(C) * H, hipMemcpyDeviceToHost));
    for (int ky = 0; ky < NY; ++ky)
        for (int kx = 0; kx < NX; ++kx) {
            double a, b;
            if (kx < NXH) { a = hs[(size_t)kx * NY + ky].x; b = hs[(size_t)kx * NY + ky].y; }
            else { C v = hs[(size_t)(NX - kx) * NY + (NY - ky) % NY]; a = v.x; b = -(double)v.y; }
            re[(size_t)ky * NX + kx] = a;
            im[(size_t)ky * NX + kx] = b;
        }
    return HW_OK;
}

template <typename T, int NX, int NY>
int fft2_inverse(const double* re, const double* im, double* phys, double* nyq_im) {
    using C = typename ctype<T>::type;
    constexpr int NXH = NX / 2 + 1;
    constexpr size_t H = (size_t)NXH * NY, N = (size_t)NX * NY;
    std::vector<C> hs(H);
    for (int kx = 0; kx < NXH; ++kx)
        for (int ky = 0; ky < NY; ++ky) {
            hs[(size_t)kx * NY + ky].x = (T)re[(size_t)ky * NX + kx];
            hs[(size_t)kx * NY + ky].y = (T)im[(size_t)ky * NX + kx];
        }
    DevBuf<C> S, X, tx, ty;
    DevBuf<T> dp, dn;
    int rc;
    if ((rc = S.upload(hs, 0)) || (rc = X.alloc(H)) || (rc = dp.alloc(N)) || (rc = dn.alloc(NY)) ||
        (rc = tx.upload(x_twiddles<T, NX>(), 0)) || (rc = ty.upload(y_twiddles<T, NY>(), 0)))
        return rc;
    k_t_yinv<T, NX, NY><<<(NXH + YGeo<T, NY>::CC - 1) / YGeo<T, NY>::CC, YGeo<T, NY>::NT>>>(S.p, X.p, ty.p);
    k_t_xinv<T, NX, NY><<<NY / XGeo<T, NX>::RR, XGeo<T, NX>::NT>>>(X.p, dp.p, dn.p, tx.p);
    HW_LAUNCH_CHECK();
    std::vector<T> hp(N), hn(NY);
    HW_TRY(hipMemcpy(hp.data(), dp.p, sizeof(T) * N, hipMemcpyDeviceToHost));
    HW_TRY(hipMemcpy(hn.data(), dn.p, sizeof(T) * NY, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < N; ++i) phys[i] = hp[i];
    if (nyq_im)
        for (int i = 0; i < NY; ++i) nyq_im[i] = hn[i];
    return HW_OK;
}

// ------------------------------------- zero-allocation forward transform --
// The forward transform of hw_fft2_forward on device arrays, with every
// intermediate (twiddle tables, the T-precision field, the x-pass and y-pass
// spectra) carved out of a caller-supplied region of device memory `ws`
// (the library's static scratch array): no allocation, no synchronisation,
// everything enqueued on `stream`.

// Twiddle tables on the device, the same formulas as x_twiddles / y_twiddles
// (angles in FP64, rounded once to T).
template <typename T, int NX, int NY>
__global__ void k_na_twiddles(typename ctype<T>::type* twx, typename ctype<T>::type* twy) {
    using XG = XGeo<T, NX>;
    using YG = YGeo<T, NY>;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    auto pass_entry = [](int L, int E, int q, double& a) -> bool {
        for (int I = 1; I < rs_np(L, E); ++I) {
            const int R = rs_radix(L, E, I), NS = rs_ns(L, E, I), o = rs_twoff(L, E, I);
            if (q >= o && q < o + (R - 1) * NS) {
                const int r = (q - o) / NS + 1, k = (q - o) % NS;
                const long m = (long)r * k * (L / (NS * R)) % L;
                a = -2.0 * M_PI * (double)m / (double)L;
                return true;
            }
        }
        return false;
    };
    if (i < XG::TWN) {
        double a = 0.0;
        bool set = i < XG::TWPP ? pass_entry(XG::L, XG::E, i, a) : true;
        if (i >= XG::TWPP) a = -2.0 * M_PI * (double)(i - XG::TWPP) / (double)NX;
        twx[i].x = set ? (T)cos(a) : (T)0;
        twx[i].y = set ? (T)sin(a) : (T)0;
    }
    if (i < YG::TWN) {
        double a = 0.0;
        bool set = pass_entry(NY, YG::E, i, a);
        twy[i].x = set ? (T)cos(a) : (T)0;
        twy[i].y = set ? (T)sin(a) : (T)0;
    }
}

// half spectrum [kx][ky] (kx = 0..NX/2) -> full-grid planes [ky][kx], the
// columns kx > NX/2 as the conjugate mirror (what hw_fft2_forward returns)
template <typename T, int NX, int NY>
__global__ void k_na_expand(const typename ctype<T>::type* __restrict__ S, double* __restrict__ re,
                            double* __restrict__ im) {
    constexpr int NXH = NX / 2 + 1;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)NX * NY) return;
    const int ky = (int)(q / NX), kx = (int)(q % NX);
    double a, b;
    if (kx < NXH) {
        const auto v = S[(size_t)kx * NY + ky];
        a = (double)v.x; b = (double)v.y;
    } else {
        const auto v = S[(size_t)(NX - kx) * NY + (NY - ky) % NY];
        a = (double)v.x; b = -(double)v.y;
    }
    re[q] = a;
    im[q] = b;
}

constexpr size_t na_align(size_t b) { return (b + 255) & ~(size_t)255; }

// bytes of scratch the zero-allocation forward transform needs at (T, NX, NY)
template <typename T, int NX, int NY> constexpr size_t fft2_na_bytes() {
    using C = typename ctype<T>::type;
    constexpr size_t H = (size_t)(NX / 2 + 1) * NY, N = (size_t)NX * NY;
    return na_align(sizeof(C) * XGeo<T, NX>::TWN) + na_align(sizeof(C) * YGeo<T, NY>::TWN) +
           na_align(sizeof(T) * N) + 2 * na_align(sizeof(C) * H);
}

template <typename T, int NX, int NY>
int fft2_forward_noalloc(const double* d_phys, double* d_re, double* d_im, hipStream_t stream, unsigned char* ws,
                         size_t ws_bytes) {
    using C = typename ctype<T>::type;
    constexpr int NXH = NX / 2 + 1;
    constexpr size_t H = (size_t)NXH * NY, N = (size_t)NX * NY;
    constexpr size_t need = fft2_na_bytes<T, NX, NY>();
    if (need > ws_bytes) {
        set_error("hw_fft2_forward_noalloc: %dx%d (%s) needs %zu bytes of scratch, the static array holds %zu "
                  "(grids up to 512x512 fit)", NX, NY, sizeof(T) == 8 ? "FP64" : "FP32", need, ws_bytes);
        return HW_ERR_ARG;
    }
    unsigned char* p = ws;
    C* twx = (C*)p; p += na_align(sizeof(C) * XGeo<T, NX>::TWN);
    C* twy = (C*)p; p += na_align(sizeof(C) * YGeo<T, NY>::TWN);
    T* ph = (T*)p; p += na_align(sizeof(T) * N);
    C* X = (C*)p; p += na_align(sizeof(C) * H);
    C* S = (C*)p;
    constexpr int TWMAX = XGeo<T, NX>::TWN > YGeo<T, NY>::TWN ? XGeo<T, NX>::TWN : YGeo<T, NY>::TWN;
    k_na_twiddles<T, NX, NY><<<(TWMAX + 255) / 256, 256, 0, stream>>>(twx, twy);
    k_convert_scalar<double, T><<<(unsigned)((N + 255) / 256 < 4096 ? (N + 255) / 256 : 4096), 256, 0, stream>>>(
        d_phys, ph, N);
    k_t_xfwd<T, NX, NY><<<NY / XGeo<T, NX>::RR, XGeo<T, NX>::NT, 0, stream>>>(ph, X, twx);
    k_t_yfwd<T, NX, NY><<<(NXH + YGeo<T, NY>::CC - 1) / YGeo<T, NY>::CC, YGeo<T, NY>::NT, 0, stream>>>(X, S, twy);
    k_na_expand<T, NX, NY><<<(unsigned)((N + 255) / 256), 256, 0, stream>>>(S, d_re, d_im);
    HW_LAUNCH_CHECK();
    return HW_OK;
}

// ------------------------------------------------------------- registry --
struct Entry {
    int prec, nx, ny;
    Solver2D* (*make)();
    int (*fwd)(const double*, double*, double*);
    int (*inv)(const double*, const double*, double*, double*);
    int (*fwd_na)(const double*, double*, double*, hipStream_t, unsigned char*, size_t);
};
std::vector<Entry>& registry();
struct Registrar {
    Registrar(Entry e) { registry().push_back(e); }
};

}  // namespace hw

#define HW_INSTANTIATE(T, PREC, NX, NY)                                                       \
    namespace {                                                                             \
    ::hw::Solver2D* make_##PREC##_##NX##_##NY() { return new ::hw::Solver2DT<T, NX, NY>(); } \
    int fwd_##PREC##_##NX##_##NY(const double* p, double* r, double* i) {                   \
        return ::hw::fft2_forward<T, NX, NY>(p, r, i);                                      \
    }                                                                                       \
    int inv_##PREC##_##NX##_##NY(const double* r, const double* i, double* p, double* n) {  \
        return ::hw::fft2_inverse<T, NX, NY>(r, i, p, n);                                   \
    }                                                                                       \
    ::hw::Registrar reg_##PREC##_##NX##_##NY({PREC, NX, NY, make_##PREC##_##NX##_##NY,       \
                                              fwd_##PREC##_##NX##_##NY, inv_##PREC##_##NX##_##NY, \
                                              ::hw::fft2_forward_noalloc<T, NX, NY>});       \
    }
