// solver1d.h -- host side of the 1-D path: a batch of independent
// realizations (uneven n), their device buffers, the fused RHS launches and
// the device-resident RKF45 driver with per-realization step control.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/hawassi_hip.h"
#include "rhs1d.h"
#include "solver2d.h"

namespace hw {

inline void natural_spline_c(const double* x, const double* yv, int n, double* cc) {
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

struct Solver1D {
    int prec = 0, nreal = 0, maxn = 0, any_brk = 0;
    long round_limit = 0;   // rounds per integrate call before HW_ERR_STEP (0: no limit)
    hipStream_t stream = nullptr;
    std::vector<Real1D> hreal;
    long long ntot = 0;
    virtual ~Solver1D() {
        if (stream) (void)hipStreamDestroy(stream);
    }
    virtual int init(const hw_config_1d* cfg, int n) = 0;
    virtual int rhs_at(int r, double t, const double* y, double* dydt) = 0;
    virtual int set_state(int r, const double* y) = 0;
    virtual int get_state(int r, double* y) = 0;
    virtual int set_time(int r, double t, double h) = 0;
    virtual int get_time(int r, double* t, double* h) = 0;
    virtual int integrate(double t1, long max_rounds) = 0;
    virtual int breaking_state(int r, double* B, int* flag) = 0;
    virtual int stats(int r, hw_stats* st) = 0;
};

// the RHS kernel instances (one per size class n) and the step-control
// kernels (one per thread count)
template <typename T> struct K1D {
    static hipError_t launch(int n, dim3 grid, size_t lds, hipStream_t s, const Batch1DArgs<T>& a, int stage, int list0,
                             int mid);
    static hipError_t control(int NT, dim3 grid, hipStream_t s, const Batch1DArgs<T>& a, int list0);
    static hipError_t set_lds(int n, size_t bytes);
};

// A size class: the realizations of one n (and one operator kind, two or
// three reference depths), consecutive in the launch list, advanced by
// their own kernel on their own stream.
struct Class1D {
    int n = 0, mid = 0, nt = 0, list0 = 0, count = 0;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr, pre = nullptr, t_a = nullptr, t_b = nullptr;
};

template <typename T> struct Solver1DT : Solver1D {
    using C = typename ctype<T>::type;
    DevBuf<Real1D> d_real;
    DevBuf<Rk1D> d_rk;
    DevBuf<int> d_list, d_count;
    DevBuf<T> k, aal, chi, damp, opa, opb, gm, gp, depth, sr, B;
    DevBuf<T> opc, gc;   // third reference depth (zero where unused)
    DevBuf<C> gamx, gamxs, gamws, Kb[7], sc;
    DevBuf<double> wst;
    DevBuf<double> tsig, sig, sigc, skw, skwc;
    DevBuf<double2> y, ext;
    DevBuf<int> peaks, ide;
    DevBuf<Cb1> cb, cbn, cbp;
    DevBuf<Brk1D> bs;
    DevBuf<C> tw;
    Batch1DArgs<T> args{};
    std::vector<Class1D> cls;
    std::vector<int> pos;          // position of realization r in the launch list
    int* h_count = nullptr;
    // each round forks from ev0 on the main stream; the third stage of the
    // first round of every integrate call (all realizations active) is timed
    // over the classes (Class1D::t_a, t_b) as one right-hand side of the
    // whole batch
    hipEvent_t ev0 = nullptr;
    hipGraph_t round_graph = nullptr;
    hipGraphExec_t round_exec = nullptr;
    double launch_ms = 0.0;
    long launches_timed = 0, rounds_total = 0;

    ~Solver1DT() override {
        if (round_exec) (void)hipGraphExecDestroy(round_exec);
        if (round_graph) (void)hipGraphDestroy(round_graph);
        if (h_count) (void)hipHostFree(h_count);
        if (ev0) (void)hipEventDestroy(ev0);
        for (Class1D& c : cls) {
            for (hipEvent_t e : {c.done, c.pre, c.t_a, c.t_b})
                if (e) (void)hipEventDestroy(e);
            if (c.stream) (void)hipStreamDestroy(c.stream);
        }
    }


    int init(const hw_config_1d* cfg, int n) override {
        nreal = n;
        prec = sizeof(T) == 4 ? HW_PREC_FP32 : HW_PREC_FP64;
        const int nmax_prec = sizeof(T) == 4 ? 16384 : 8192;
        long long cbtot = 0, sigtot = 0;
        hreal.resize(n);
        for (int r = 0; r < n; ++r) {
            const hw_config_1d& c = cfg[r];
            int nn = c.n;
            // the sizes with a kernel: powers of two, and 2000 (the grid of the
            // reference case 1dBBr001HarmBar: radix 8, 2, 5, 5, 5 passes)
            if (!((nn >= 128 && nn <= nmax_prec && !(nn & (nn - 1))) || nn == 2000)) {
                set_error("realization %d: n = %d unsupported (powers of two 128..%d, or 2000)", r, nn, nmax_prec);
                return HW_ERR_ARG;
            }
            if (!c.k || !c.aal || !c.chi || !c.damp || !c.depth) { set_error("realization %d: null array", r); return HW_ERR_ARG; }
            if (c.flat ? (!c.csqr || !c.lop) : (!c.csq_min || !c.csq_plus || !c.gam_min || !c.gam_plus)) {
                set_error("realization %d: operator arrays missing", r);
                return HW_ERR_ARG;
            }
            if (c.source_on && (c.nt < 2 || !c.t_sig || !c.sig || !c.sig_skew || !c.gamx_re || !c.gamx_im ||
                                !c.gamx_skew_re || !c.gamx_skew_im)) {
                set_error("realization %d: influx arrays missing", r);
                return HW_ERR_ARG;
            }
            Real1D& R = hreal[r];
            std::memset(&R, 0, sizeof(R));
            R.n = nn; R.flat = c.flat; R.src = c.source_on; R.prop = c.propagation; R.brk = c.breaking_on;
            R.nt = c.source_on ? c.nt : 0;
            R.cap = nn / 4 < 1024 ? nn / 4 : 1024;
            R.off = ntot;
            R.sig_off = sigtot;
            R.cb_off = cbtot;
            R.dx = c.dx; R.KBC = c.KBC; R.TC = c.TC; R.Tstar = c.Tstar; R.MaxEtaInit = c.max_eta_init;
            R.delb = c.delb; R.eps_abs = c.eps_abs; R.eps_rel = c.eps_rel;
            R.xbstart = c.xbstart; R.xbend = c.xbend; R.xbendL = c.xbendL; R.xbstartR = c.xbstartR;
            R.lenIF = c.lenIF; R.ILL = c.ILL; R.IFF = c.IFF;
            R.mid = !c.flat && c.gam_mid != nullptr;
            R.order = c.order == 0 ? 2 : c.order;
            if (R.order < 1 || R.order > 3) { set_error("realization %d: order must be 1, 2 or 3", r); return HW_ERR_ARG; }
            if (R.order == 1) R.brk = 0;   // rhs_HsF1 never calls rhs_breaking
            if (R.mid && !c.csq_mid) { set_error("realization %d: gam_mid without csq_mid", r); return HW_ERR_ARG; }
            // rhs_HsF3 neither updates nor adds the wall source (1d/rhs.c:243-399)
            R.wall = c.wall_on && R.order != 3;
            if (R.wall) {
                if (!c.gamw_skew_re || !c.gamw_skew_im || c.idx_wall < 0 || c.idx_wall >= nn) {
                    set_error("realization %d: wall node outside the grid or kernel missing", r);
                    return HW_ERR_ARG;
                }
                R.idx_wall = c.idx_wall;
                R.refl = c.refl_coef;
            }
            if (R.brk) {
                int hi = R.xbstart + (R.xbendL == 0 && R.xbstartR == 0 ? R.lenIF : R.ILL);
                int hi2 = R.xbstartR + R.IFF;
                if (R.xbstart < 0 || R.xbend >= nn || hi > nn || hi2 > nn || R.xbstartR < 0 || c.dx <= 0) {
                    set_error("realization %d: breaking window outside the grid", r);
                    return HW_ERR_ARG;
                }
            }
            ntot += nn;
            sigtot += R.nt;
            cbtot += R.cap;
            if (nn > maxn) maxn = nn;
            // the stepper's dydt_out evaluation has side effects only through
            // breaking and the wall source
            any_brk |= R.brk | R.wall;
        }
        HW_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        // ---- size classes: the launch list holds each class's realizations
        // together (in realization order within a class)
        std::vector<int> hl;
        pos.assign(n, 0);
        for (int r = 0; r < n; ++r) {
            bool seen = false;
            for (const Class1D& c : cls) seen |= c.n == hreal[r].n && c.mid == hreal[r].mid;
            if (seen) continue;
            Class1D c;
            c.n = hreal[r].n;
            c.mid = hreal[r].mid;
            c.nt = nt_for_n(c.n);
            c.list0 = (int)hl.size();
            for (int q = r; q < n; ++q)
                if (hreal[q].n == c.n && hreal[q].mid == c.mid) {
                    pos[q] = (int)hl.size();
                    hl.push_back(q);
                }
            c.count = (int)hl.size() - c.list0;
            cls.push_back(c);
        }
        for (Class1D& c : cls) {
            HW_TRY(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
            HW_TRY(hipEventCreateWithFlags(&c.done, hipEventDisableTiming));
            HW_TRY(hipEventCreateWithFlags(&c.pre, hipEventDisableTiming));
            HW_TRY(hipEventCreate(&c.t_a));
            HW_TRY(hipEventCreate(&c.t_b));
        }
        // ---- per-element arrays
        std::vector<T> hk(ntot), haal(ntot), hchi(ntot), hdamp(ntot), hopa(ntot), hopb(ntot), hgm(ntot), hgp(ntot),
            hdep(ntot), hopc(ntot, (T)0), hgc(ntot, (T)0);
        std::vector<C> hgx(ntot), hgs(ntot), hgw(ntot);
        std::vector<double> hts(sigtot), hsig(sigtot), hsigc(sigtot), hskw(sigtot), hskwc(sigtot);
        for (int r = 0; r < n; ++r) {
            const hw_config_1d& c = cfg[r];
            const Real1D& R = hreal[r];
            for (int i = 0; i < R.n; ++i) {
                size_t q = R.off + i;
                hk[q] = (T)c.k[i]; haal[q] = (T)c.aal[i]; hchi[q] = (T)c.chi[i]; hdamp[q] = (T)c.damp[i];
                hdep[q] = (T)c.depth[i];
                if (c.flat) { hopa[q] = (T)c.csqr[i]; hopb[q] = (T)c.lop[i]; hgm[q] = 0; hgp[q] = 0; }
                else {
                    hopa[q] = (T)c.csq_min[i]; hopb[q] = (T)c.csq_plus[i];
                    hgm[q] = (T)c.gam_min[i]; hgp[q] = (T)c.gam_plus[i];
                    if (R.mid) { hopc[q] = (T)c.csq_mid[i]; hgc[q] = (T)c.gam_mid[i]; }
                }
                if (c.source_on) {
                    hgx[q].x = (T)c.gamx_re[i]; hgx[q].y = (T)c.gamx_im[i];
                    hgs[q].x = (T)c.gamx_skew_re[i]; hgs[q].y = (T)c.gamx_skew_im[i];
                } else {
                    hgx[q].x = hgx[q].y = hgs[q].x = hgs[q].y = (T)0;
                }
                hgw[q].x = R.wall ? (T)c.gamw_skew_re[i] : (T)0;
                hgw[q].y = R.wall ? (T)c.gamw_skew_im[i] : (T)0;
            }
            if (R.nt) {
                std::memcpy(&hts[R.sig_off], c.t_sig, sizeof(double) * R.nt);
                std::memcpy(&hsig[R.sig_off], c.sig, sizeof(double) * R.nt);
                std::memcpy(&hskw[R.sig_off], c.sig_skew, sizeof(double) * R.nt);
                natural_spline_c(c.t_sig, c.sig, R.nt, &hsigc[R.sig_off]);
                natural_spline_c(c.t_sig, c.sig_skew, R.nt, &hskwc[R.sig_off]);
            }
        }
        int rc;
        if ((rc = k.upload(hk, stream)) || (rc = aal.upload(haal, stream)) || (rc = chi.upload(hchi, stream)) ||
            (rc = damp.upload(hdamp, stream)) || (rc = opa.upload(hopa, stream)) || (rc = opb.upload(hopb, stream)) ||
            (rc = gm.upload(hgm, stream)) || (rc = gp.upload(hgp, stream)) || (rc = depth.upload(hdep, stream)) ||
            (rc = gamx.upload(hgx, stream)) || (rc = gamxs.upload(hgs, stream)) || (rc = gamws.upload(hgw, stream)) ||
            (rc = opc.upload(hopc, stream)) ||
            (rc = gc.upload(hgc, stream)))
            return rc;
        if (sigtot) {
            if ((rc = tsig.upload(hts, stream)) || (rc = sig.upload(hsig, stream)) || (rc = sigc.upload(hsigc, stream)) ||
                (rc = skw.upload(hskw, stream)) || (rc = skwc.upload(hskwc, stream)))
                return rc;
        }
        // ---- twiddles exp(-2 pi i m / n) for every size present, one table
        // per size in one buffer (a realization finds its table at tw_off)
        {
            std::vector<C> h;
            std::vector<std::pair<int, long long>> at;   // (n, offset of its table)
            for (int r = 0; r < n; ++r) {
                const int nn = hreal[r].n;
                long long off = -1;
                for (const auto& e : at)
                    if (e.first == nn) off = e.second;
                if (off < 0) {
                    off = (long long)h.size();
                    at.emplace_back(nn, off);
                    for (int m = 0; m < nn; ++m) {
                        const double ang = -2.0 * M_PI * (double)m / (double)nn;
                        C w;
                        w.x = (T)std::cos(ang);
                        w.y = (T)std::sin(ang);
                        h.push_back(w);
                    }
                }
                hreal[r].tw_off = off;
            }
            if ((rc = tw.upload(h, stream))) return rc;
        }
        // ---- state, stages, scratch, breaking
        std::vector<Rk1D> hrk(n);
        for (int r = 0; r < n; ++r) {
            std::memset(&hrk[r], 0, sizeof(Rk1D));
            hrk[r].t = cfg[r].t0;
            hrk[r].h = cfg[r].h0 > 0 ? cfg[r].h0 : 1e-4;
        }
        std::vector<Brk1D> hb(n);
        std::memset(hb.data(), 0, sizeof(Brk1D) * n);
        // wall source state: t_prev = the first record time (define_wall,
        // 1d/functions.c:9), S = 0, previous refl*eta = 0, calls = 0
        std::vector<double> hw(4 * (size_t)n, 0.0);
        for (int r = 0; r < n; ++r) hw[4 * r] = cfg[r].t_wall0;
        if ((rc = d_real.upload(hreal, stream)) || (rc = d_rk.upload(hrk, stream)) || (rc = d_list.upload(hl, stream)) ||
            (rc = d_count.alloc(2)) || (rc = y.alloc(2 * ntot)) || (rc = ext.alloc(2 * (size_t)maxn)) ||
            (rc = sr.alloc(R1_NR * ntot)) || (rc = sc.alloc(C1_NC * ntot)) || (rc = B.alloc(ntot)) ||
            (rc = peaks.alloc(ntot)) || (rc = ide.alloc(cbtot)) || (rc = cb.alloc(cbtot)) || (rc = cbn.alloc(cbtot)) ||
            (rc = cbp.alloc(cbtot)) || (rc = bs.upload(hb, stream)) || (rc = wst.upload(hw, stream)))
            return rc;
        for (int j = 0; j < 7; ++j)
            if ((rc = Kb[j].alloc(2 * ntot))) return rc;
        HW_TRY(hipMemsetAsync(y.p, 0, sizeof(double2) * 2 * ntot, stream));
        HW_TRY(hipMemsetAsync(B.p, 0, sizeof(T) * ntot, stream));
        HW_TRY(hipMemsetAsync(cb.p, 0, sizeof(Cb1) * cbtot, stream));
        HW_TRY(hipMemsetAsync(cbn.p, 0, sizeof(Cb1) * cbtot, stream));
        HW_TRY(hipMemsetAsync(cbp.p, 0, sizeof(Cb1) * cbtot, stream));
        HW_TRY(hipMemsetAsync(ide.p, 0, sizeof(int) * cbtot, stream));
        HW_TRY(hipHostMalloc((void**)&h_count, 2 * sizeof(int), hipHostMallocDefault));
        HW_TRY(hipEventCreateWithFlags(&ev0, hipEventDisableTiming));
        for (const Class1D& c : cls) HW_TRY(K1D<T>::set_lds(c.n, sizeof(C) * c.n));
        // ---- kernel arguments
        Batch1DArgs<T>& a = args;
        a.real = d_real.p; a.rk = d_rk.p; a.list = d_list.p;
        a.k = k.p; a.aal = aal.p; a.chi = chi.p; a.damp = damp.p; a.opa = opa.p; a.opb = opb.p;
        a.gm = gm.p; a.gp = gp.p; a.depth = depth.p; a.gamx = gamx.p; a.gamxs = gamxs.p;
        a.gamws = gamws.p; a.wst = wst.p;
        a.opc = opc.p; a.gc = gc.p;
        a.tsig = tsig.p; a.sig = sig.p; a.sigc = sigc.p; a.skw = skw.p; a.skwc = skwc.p;
        a.y = y.p;
        for (int j = 0; j < 7; ++j) a.K[j] = Kb[j].p;
        a.ext = ext.p;
        a.sr = sr.p; a.sc = sc.p; a.B = B.p; a.peaks = peaks.p; a.cb = cb.p; a.cbn = cbn.p; a.cbp = cbp.p;
        a.ide = ide.p; a.bs = bs.p;
        a.tw = tw.p;
        a.g = (T)9.80665;
        HW_TRY(hipStreamSynchronize(stream));
        return HW_OK;
    }

    int check_r(int r) {
        if (r < 0 || r >= nreal) { set_error("realization %d out of range (batch of %d)", r, nreal); return HW_ERR_ARG; }
        return HW_OK;
    }

    // 4-plane host layout <-> [eta^][u^] double2
    int rhs_at(int r, double t, const double* yin, double* dydt) override {
        int rc = check_r(r);
        if (rc) return rc;
        const int n = hreal[r].n;
        std::vector<double2> h(2 * n);
        for (int i = 0; i < n; ++i) {
            h[i] = make_double2(yin[i], yin[n + i]);
            h[n + i] = make_double2(yin[2 * n + i], yin[3 * n + i]);
        }
        HW_TRY(hipMemcpyAsync(ext.p + 0, h.data(), sizeof(double2) * 2 * n, hipMemcpyHostToDevice, stream));
        // the stand-alone input sits at the realization's own offset in ext
        Batch1DArgs<T> a = args;
        a.ext = ext.p - 2 * hreal[r].off;
        a.t_ext = t;
        HW_TRY(K1D<T>::launch(n, dim3(1), sizeof(C) * n, stream, a, 0, pos[r], hreal[r].mid));
        std::vector<C> o(2 * n);
        HW_TRY(hipMemcpyAsync(o.data(), Kb[6].p + 2 * hreal[r].off, sizeof(C) * 2 * n, hipMemcpyDeviceToHost, stream));
        HW_TRY(hipStreamSynchronize(stream));
        for (int i = 0; i < n; ++i) {
            dydt[i] = (double)o[i].x; dydt[n + i] = (double)o[i].y;
            dydt[2 * n + i] = (double)o[n + i].x; dydt[3 * n + i] = (double)o[n + i].y;
        }
        return HW_OK;
    }

    int set_state(int r, const double* yin) override {
        int rc = check_r(r);
        if (rc) return rc;
        const int n = hreal[r].n;
        std::vector<double2> h(2 * n);
        for (int i = 0; i < n; ++i) {
            h[i] = make_double2(yin[i], yin[n + i]);
            h[n + i] = make_double2(yin[2 * n + i], yin[3 * n + i]);
        }
        HW_TRY(hipMemcpyAsync(y.p + 2 * hreal[r].off, h.data(), sizeof(double2) * 2 * n, hipMemcpyHostToDevice, stream));
        HW_TRY(hipStreamSynchronize(stream));
        return HW_OK;
    }

    int get_state(int r, double* yout) override {
        int rc = check_r(r);
        if (rc) return rc;
        const int n = hreal[r].n;
        std::vector<double2> h(2 * n);
        HW_TRY(hipMemcpyAsync(h.data(), y.p + 2 * hreal[r].off, sizeof(double2) * 2 * n, hipMemcpyDeviceToHost, stream));
        HW_TRY(hipStreamSynchronize(stream));
        for (int i = 0; i < n; ++i) {
            yout[i] = h[i].x; yout[n + i] = h[i].y;
            yout[2 * n + i] = h[n + i].x; yout[3 * n + i] = h[n + i].y;
        }
        return HW_OK;
    }

    int set_time(int r, double t, double h) override {
        int rc = check_r(r);
        if (rc) return rc;
        Rk1D s;
        HW_TRY(hipMemcpyAsync(&s, d_rk.p + r, sizeof(Rk1D), hipMemcpyDeviceToHost, stream));
        HW_TRY(hipStreamSynchronize(stream));
        s.t = t;
        s.h = h;
        HW_TRY(hipMemcpyAsync(d_rk.p + r, &s, sizeof(Rk1D), hipMemcpyHostToDevice, stream));
        HW_TRY(hipStreamSynchronize(stream));
        return HW_OK;
    }

    int get_time(int r, double* t, double* h) override {
        int rc = check_r(r);
        if (rc) return rc;
        Rk1D s;
        HW_TRY(hipMemcpyAsync(&s, d_rk.p + r, sizeof(Rk1D), hipMemcpyDeviceToHost, stream));
        HW_TRY(hipStreamSynchronize(stream));
        if (t) *t = s.t;
        if (h) *h = s.h;
        return HW_OK;
    }

    // One round of every active realization, enqueued: the classes fork from
    // the main stream, run their stages (dydt_in where needed, five stages,
    // rkf45's dydt_out when breaking or a wall is on) and their step control
    // on their own streams, and join it again for the active count.  With
    // `timed` the classes meet before the third stage, whose launches side by
    // side are one right-hand side of the whole batch, timed per class.
    int enqueue_round(bool timed) {
        HW_TRY(hipEventRecord(ev0, stream));
        for (Class1D& c : cls) HW_TRY(hipStreamWaitEvent(c.stream, ev0, 0));
        for (int s = 1; s <= 6; ++s) {
            if (timed && s == 3 && cls.size() > 1) {
                for (Class1D& c : cls) HW_TRY(hipEventRecord(c.pre, c.stream));
                for (Class1D& c : cls)
                    for (Class1D& d : cls)
                        if (&c != &d) HW_TRY(hipStreamWaitEvent(c.stream, d.pre, 0));
            }
            for (Class1D& c : cls) {
                if (timed && s == 3) HW_TRY(hipEventRecord(c.t_a, c.stream));
                HW_TRY(K1D<T>::launch(c.n, dim3(c.count), sizeof(C) * c.n, c.stream, args, s, c.list0, c.mid));
                if (timed && s == 3) HW_TRY(hipEventRecord(c.t_b, c.stream));
            }
        }
        for (Class1D& c : cls) {
            if (any_brk)
                HW_TRY(K1D<T>::launch(c.n, dim3(c.count), sizeof(C) * c.n, c.stream, args, 7, c.list0, c.mid));
            HW_TRY(K1D<T>::control(c.nt, dim3(c.count), c.stream, args, c.list0));
            HW_TRY(hipEventRecord(c.done, c.stream));
            HW_TRY(hipStreamWaitEvent(stream, c.done, 0));
        }
        k1d_count<<<1, 256, 0, stream>>>(d_rk.p, bs.p, nreal, d_count.p);
        HW_LAUNCH_CHECK();
        return HW_OK;
    }

    // the untimed round captured once into a HIP graph (fork and join
    // included); replays of it advance every active realization by one
    // attempted step, and a replay after every realization has reached t1
    // only launches blocks that return at once
    int capture_round() {
        HW_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_round(false);
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamEndCapture(stream, &gr);
        if (rc) {
            if (gr) (void)hipGraphDestroy(gr);
            return rc;
        }
        HW_TRY(e);
        round_graph = gr;
        HW_TRY(hipGraphInstantiate(&round_exec, round_graph, nullptr, nullptr, 0));
        return HW_OK;
    }

    // every realization: while (t < t1) gsl_odeiv_evolve_apply (ode_solver_Ham's
    // inner loop, 1d/functions.c:546-552).  Each round is one attempt of
    // every active realization.  The first round of a call runs eagerly and
    // is timed (every realization is active then); the later rounds are
    // replays of the captured round, kGraphRounds between two reads of the
    // active count.
    static constexpr long kGraphRounds = 4;
    int integrate(double t1, long max_rounds) override {
        k1d_begin<<<(nreal + 255) / 256, 256, 0, stream>>>(d_rk.p, nreal, t1);
        HW_LAUNCH_CHECK();
        if (round_limit > 0 && (max_rounds <= 0 || round_limit < max_rounds)) max_rounds = round_limit;
        int rc;
        for (long round = 0; max_rounds <= 0 || round < max_rounds;) {
            const bool timed = round == 0;
            long k = 1;
            if (timed) {
                if ((rc = enqueue_round(true))) return rc;
            } else {
                if (!round_exec && (rc = capture_round())) return rc;
                k = kGraphRounds;
                if (max_rounds > 0 && max_rounds - round < k) k = max_rounds - round;
                for (long i = 0; i < k; ++i) HW_TRY(hipGraphLaunch(round_exec, stream));
            }
            round += k;
            rounds_total += k;
            HW_TRY(hipMemcpyAsync(h_count, d_count.p, 2 * sizeof(int), hipMemcpyDeviceToHost, stream));
            HW_TRY(hipStreamSynchronize(stream));
            if (timed) {
                // the classes run the third stage side by side from the same
                // start; the batch's stage ends with its slowest class
                float ms_max = 0.f;
                for (const Class1D& c : cls) {
                    float ms = 0.f;
                    HW_TRY(hipEventElapsedTime(&ms, c.t_a, c.t_b));
                    if (ms > ms_max) ms_max = ms;
                }
                launch_ms += ms_max;
                launches_timed += 1;
            }
            if (h_count[1] > 0) return check_overflow();
            if (h_count[0] == 0) return check_overflow();
        }
        set_error("integrate: %ld rounds without reaching t1 = %g", max_rounds, t1);
        return HW_ERR_STEP;
    }

    int check_overflow() {
        std::vector<Brk1D> hb(nreal);
        HW_TRY(hipMemcpyAsync(hb.data(), bs.p, sizeof(Brk1D) * nreal, hipMemcpyDeviceToHost, stream));
        HW_TRY(hipStreamSynchronize(stream));
        for (int r = 0; r < nreal; ++r)
            if (hb[r].overflow) {
                set_error("realization %d: more breaking crests than the crest list holds (%d)", r, hreal[r].cap);
                return HW_ERR_ARG;
            }
        return HW_OK;
    }

    int breaking_state(int r, double* Bout, int* flag) override {
        int rc = check_r(r);
        if (rc) return rc;
        const int n = hreal[r].n;
        if (Bout) {
            std::vector<T> h(n);
            HW_TRY(hipMemcpyAsync(h.data(), B.p + hreal[r].off, sizeof(T) * n, hipMemcpyDeviceToHost, stream));
            HW_TRY(hipStreamSynchronize(stream));
            for (int i = 0; i < n; ++i) Bout[i] = (double)h[i];
        }
        if (flag) {
            Brk1D b;
            HW_TRY(hipMemcpyAsync(&b, bs.p + r, sizeof(Brk1D), hipMemcpyDeviceToHost, stream));
            HW_TRY(hipStreamSynchronize(stream));
            *flag = b.flag;
        }
        return check_overflow();
    }

    int stats(int r, hw_stats* st) override {
        std::vector<Rk1D> h(nreal);
        HW_TRY(hipMemcpyAsync(h.data(), d_rk.p, sizeof(Rk1D) * nreal, hipMemcpyDeviceToHost, stream));
        HW_TRY(hipStreamSynchronize(stream));
        *st = hw_stats{};
        for (int q = 0; q < nreal; ++q) {
            if (r >= 0 && q != r) continue;
            st->rhs_evals += h[q].rhs;
            st->steps_accepted += h[q].accepted;
            st->steps_rejected += h[q].rejected;
        }
        if (r < 0) {          // per launch, all realizations of the batch together
            st->rhs_ms_total = launch_ms;
            st->rhs_timed = launches_timed;
        }
        return HW_OK;
    }
};

}  // namespace hw
