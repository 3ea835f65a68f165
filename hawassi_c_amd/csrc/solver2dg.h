// solver2dg.h -- the full-grid 2-D path: the reference's right-hand side on
// its own full complex grid, for any grid whose sides factor into 2, 3 and 5.
//
// The fused half-spectrum path (solver2d.h) covers power-of-two grids and
// advances only the half of the spectrum that feeds the dynamics.  This path
// keeps what the reference keeps (2d/fun_odesolver.c:1152-1160): the ODE
// state is the full 6-plane complex grid, every transform is a full complex
// 2-D FFT (fftg.h), the Hermitian projection fftwcomplex_2d_sym
// (2d/functions1.c:492-521) is applied where the reference applies it, and
// the output is formed on every point, so the columns kx > nx/2 the
// reference's epilogue writes (2d/fun_odesolver.c:4589-4613) are its values,
// not a mirror, and the step control's max-norm runs over all 6 nx ny values
// as gsl_odeiv2_evolve_apply's does.  It also has the flat-bottom wall
// branches of fun_delphiH1hat (2d/fun_odesolver.c:2100-2239).
//
// The evaluation follows the reference's sequence (rhs_AB{1,2,3}_2D{F,B}_uv,
// fun_termshat_*, HSSgen_2depth/_3depth, fun_adjust, fun_source,
// breaking_process + fun_sbreaking), with every pointwise loop one kernel
// (k_map with a device lambda) and transforms in the precision T.  It is
// chosen for grids without compiled fused kernels and when hw_config_2d
// asks for the full grid.
#pragma once
#include "fftg.h"
#include "solver2d.h"

namespace hw {

template <typename T> struct Solver2DG : Solver2D {
    using C = typename ctype<T>::type;
    size_t N = 0;
    Fft2G<C> fft;
    DevBuf<T> kxv, kyv, aal, fbdy7, wall, chi, c2d, l2d, gwx, gwy;
    DevBuf<T> gam_m, gam_p, gam_c, csq_m, csq_p, csq_c, om2_m, om2_p, om2_c;
    DevBuf<T> newgam, skewgam, src_sig, src_skew;
    DevBuf<double> sig_y, sig_c, skew_y, skew_c;
    std::vector<double> tsig;
    int nline = 0, nt = 0, src_on = 0, vertical = 0, propagation = 0;
    bool flat = false, hs1 = false, hs3 = false, three = false, wall_on = false, refl_one = false;
    double g = 9.80665, eps_abs = 1e-3, eps_rel = 1e-2;
    DevBuf<double2> y, ynew;
    DevBuf<C> kbuf, ytmp;
    DevBuf<unsigned long long> rmax;
    unsigned long long* h_rmax = nullptr;
    DevBuf<C> pool;                  // NBUF work arrays of N
    static constexpr int NBUF = 56;
    C* W[NBUF] = {};
    // breaking
    int brk_on = 0;
    double delb = 1.5;
    CrestMachine<T> cm;
    DevBuf<T> ep, up, vp, hil, flux, sbxp, sbyp;
    DevBuf<C> Sbx, Sby;              // Sb_x^, Sb_y^ full spectra (kept between calls)
    bool last_gated = false;
    // every fourth evaluation bracketed by events
    static constexpr int NEV = 64;
    hipEvent_t ev_a[NEV] = {}, ev_b[NEV] = {};
    int nev = 0;

    ~Solver2DG() override {
        if (h_rmax) (void)hipHostFree(h_rmax);
        for (int i = 0; i < NEV; ++i) {
            if (ev_a[i]) (void)hipEventDestroy(ev_a[i]);
            if (ev_b[i]) (void)hipEventDestroy(ev_b[i]);
        }
        if (d_full_a) (void)hipFree(d_full_a);
        if (d_full_b) (void)hipFree(d_full_b);
        if (stream) (void)hipStreamDestroy(stream);
    }

    static dim3 grid1(size_t n) {
        size_t b = (n + 255) / 256;
        if (b > 8192) b = 8192;
        return dim3((unsigned)(b ? b : 1));
    }
    template <class F> void map(size_t n, F f) { k_map<<<grid1(n), 256, 0, stream>>>(n, f); }
    void proj(C* a) { k_proj_full<C><<<grid1(N), 256, 0, stream>>>(a, nx, ny); }
    void fwd(C* a) { fft.run(a, -1, stream); }
    void bwd(C* a) { fft.run(a, +1, stream); }
    void copy(C* d, const C* s) { (void)hipMemcpyAsync(d, s, sizeof(C) * N, hipMemcpyDeviceToDevice, stream); }

    std::vector<T> to_t(const double* a, size_t n, double scale = 1.0) {
        std::vector<T> o(n);
        for (size_t i = 0; i < n; ++i) o[i] = (T)(a[i] * scale);
        return o;
    }
    int up_field(DevBuf<T>& b, const double* a, double scale = 1.0) { return b.upload(to_t(a, N, scale), stream); }

    int init(const hw_config_2d* c) override {
        nx = c->nx; ny = c->ny; prec = c->precision;
        N = (size_t)nx * ny;
        if (nb != 1) { set_error("the full-grid path holds one state per context"); return HW_ERR_ARG; }
        if (c->model < 0 || c->model > 3) {
            set_error("hw_config_2d.model must be 0 or 2 (HS2), 1 (HS1) or 3 (HS3), got %d", c->model);
            return HW_ERR_ARG;
        }
        flat = c->flat != 0;
        hs1 = c->model == 1;
        hs3 = c->model == 3;
        g = c->grav; eps_abs = c->eps_abs; eps_rel = c->eps_rel;
        wall_on = c->wall != 0;
        refl_one = c->refl_one != 0;
        HW_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        for (int i = 0; i < NEV; ++i) {
            HW_TRY(hipEventCreate(&ev_a[i]));
            HW_TRY(hipEventCreate(&ev_b[i]));
        }
        if (const char* m = fft.init(nx, ny, stream)) { set_error("full-grid path, grid %dx%d: %s", nx, ny, m); return HW_ERR_ARG; }
        if (!c->kx || !c->ky || !c->aal || !c->fbdy || !c->wallchar || !c->chi) {
            set_error("hw_config_2d: kx, ky, aal, fbdy, wallchar and chi are required");
            return HW_ERR_ARG;
        }
        int rc;
        if ((rc = kxv.upload(to_t(c->kx, nx), stream)) || (rc = kyv.upload(to_t(c->ky, ny), stream)) ||
            (rc = up_field(aal, c->aal)) || (rc = up_field(fbdy7, c->fbdy, 7.0 * c->cpeak)) ||
            (rc = up_field(wall, c->wallchar)) || (rc = up_field(chi, c->chi)))
            return rc;
        if (flat) {
            if (!c->c2d || !c->l2d) { set_error("flat model needs c2d and l2d"); return HW_ERR_ARG; }
            if ((rc = up_field(c2d, c->c2d)) || (rc = up_field(l2d, c->l2d))) return rc;
            if (wall_on && refl_one) {
                if (!c->grad_wall_x || !c->grad_wall_y) {
                    set_error("flat bottom with walls of reflection 1 needs grad_wall_x and grad_wall_y");
                    return HW_ERR_ARG;
                }
                if ((rc = up_field(gwx, c->grad_wall_x)) || (rc = up_field(gwy, c->grad_wall_y))) return rc;
            }
        } else {
            if (!c->gam_min || !c->gam_plus || !c->csq_min || !c->csq_plus || !c->om2_min || !c->om2_plus) {
                set_error("two-depth model needs gam/csq/om2 min and plus");
                return HW_ERR_ARG;
            }
            if ((rc = up_field(gam_m, c->gam_min)) || (rc = up_field(gam_p, c->gam_plus)) ||
                (rc = up_field(csq_m, c->csq_min)) || (rc = up_field(csq_p, c->csq_plus)) ||
                (rc = up_field(om2_m, c->om2_min)) || (rc = up_field(om2_p, c->om2_plus)))
                return rc;
            three = c->gam_mid != nullptr;
            if (three) {
                if (!c->csq_mid || !c->om2_mid) {
                    set_error("three reference depths need gam_mid, csq_mid and om2_mid");
                    return HW_ERR_ARG;
                }
                if ((rc = up_field(gam_c, c->gam_mid)) || (rc = up_field(csq_c, c->csq_mid)) ||
                    (rc = up_field(om2_c, c->om2_mid)))
                    return rc;
            }
        }
        src_on = c->source_on;
        if (src_on) {
            nline = c->nline; nt = c->nt;
            vertical = c->vertical; propagation = c->propagation;
            int want = c->vertical ? ny : nx;
            if (nline != want || nt < 3 || !c->t_sig || !c->sig || !c->skew_sig || !c->new_gam || !c->skew_gam) {
                set_error("source: nline must be %d and t_sig/sig/skew_sig/new_gam/skew_gam given (nt>=3)", want);
                return HW_ERR_ARG;
            }
            tsig.assign(c->t_sig, c->t_sig + nt);
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
                (rc = skew_c.upload(kc, stream)) || (rc = src_sig.alloc(nline)) || (rc = src_skew.alloc(nline)) ||
                (rc = up_field(newgam, c->new_gam)) || (rc = up_field(skewgam, c->skew_gam)))
                return rc;
        }
        tprev = c->t0;
        brk_on = c->breaking_on && !hs1;   // rhs_AB1_* never calls breaking_process
        if (brk_on) {
            if ((rc = cm.init(c, nx, ny, stream))) return rc;
            delb = c->delb;
            if ((rc = hil.alloc(N)) || (rc = flux.alloc(N)) || (rc = sbxp.alloc(N)) || (rc = sbyp.alloc(N)) ||
                (rc = Sbx.alloc(N)) || (rc = Sby.alloc(N)))
                return rc;
            k_fill<C><<<grid1(N), 256, 0, stream>>>(Sbx.p, cmk_h(0, 0), N);
            k_fill<C><<<grid1(N), 256, 0, stream>>>(Sby.p, cmk_h(0, 0), N);
            HW_LAUNCH_CHECK();
        }
        if ((rc = ep.alloc(N)) || (rc = up.alloc(N)) || (rc = vp.alloc(N))) return rc;
        if ((rc = pool.alloc(NBUF * N))) return rc;
        for (int i = 0; i < NBUF; ++i) W[i] = pool.p + (size_t)i * N;
        if ((rc = y.alloc(3 * N)) || (rc = ynew.alloc(3 * N)) || (rc = kbuf.alloc(6 * 3 * N)) || (rc = ytmp.alloc(3 * N)) ||
            (rc = rmax.alloc(1)))
            return rc;
        HW_TRY(hipHostMalloc((void**)&h_rmax, sizeof(unsigned long long), hipHostMallocDefault));
        HW_TRY(hipMemsetAsync(y.p, 0, sizeof(double2) * 3 * N, stream));
        HW_TRY(hipStreamSynchronize(stream));
        return HW_OK;
    }

    static C cmk_h(T a, T b) { C r; r.x = a; r.y = b; return r; }

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

    // ------------------------------------------------------ operators --
    // HSSgen_2depth / HSSgen_3depth (2d/fun_odesolver.c:500-697):
    // acc = sum_j (F(gam_j F^-1(P Om_j phihat)/N) + Om_j F(gam_j phi / fact)) / 2
    // over the reference depths (min, mid, plus); phi holds the physical
    // field times fact in its real part.
    void hss(const C* phi, const C* phihat, const T* Om, const T* Oc, const T* Op, C* acc, double fact) {
        const size_t n = N;
        const T* Gs[3] = {gam_m.p, three ? gam_c.p : nullptr, gam_p.p};
        const T* Os[3] = {Om, three ? Oc : nullptr, Op};
        C* L = W[50];
        C* gg = W[51];
        C* gL = W[52];
        map(n, [=] __device__(size_t q) { acc[q] = cmk<C>(0, 0); });
        for (int d = 0; d < 3; ++d) {
            if (!Os[d]) continue;
            const T* O = Os[d];
            const T* G = Gs[d];
            map(n, [=] __device__(size_t q) { L[q] = cmk<C>(O[q] * phihat[q].x, O[q] * phihat[q].y); });
            proj(L);
            bwd(L);
            const T f = (T)fact, nn = (T)n;
            map(n, [=] __device__(size_t q) {
                gg[q] = cmk<C>(G[q] * phi[q].x / f, 0);
                gL[q] = cmk<C>(G[q] * L[q].x / nn, 0);
            });
            fwd(gg);
            fwd(gL);
            map(n, [=] __device__(size_t q) {
                acc[q].x += (gL[q].x + O[q] * gg[q].x) / 2;
                acc[q].y += (gL[q].y + O[q] * gg[q].y) / 2;
            });
        }
    }

    // fun_adjust (2d/fun_odesolver.c:810-833): out = F(chi F^-1(P A)/N),
    // the inverse's imaginary part kept; A is projected in place
    void adjust(C* out, C* A) {
        const T* ch = chi.p;
        const T nn = (T)N;
        proj(A);
        copy(out, A);
        bwd(out);
        map(N, [=] __device__(size_t q) { out[q] = cmk<C>(out[q].x * ch[q] / nn, out[q].y * ch[q] / nn); });
        fwd(out);
    }

    // fun_source (2d/fun_odesolver.c:886-983): Source_hat at time t
    void source_hat(double t, C* S) {
        if (!src_on) {
            map(N, [=] __device__(size_t q) { S[q] = cmk<C>(0, 0); });
            return;
        }
        int idx = bsearch_interval(tsig, t);
        double hh = tsig[idx + 1] - tsig[idx];
        k_src_eval<T><<<(nline + 255) / 256, 256, 0, stream>>>(sig_y.p, sig_c.p, skew_y.p, skew_c.p, nline, idx,
                                                               t - tsig[idx], hh, src_sig.p, src_skew.p);
        C* Sk = W[53];
        const T *ng = newgam.p, *sg = skewgam.p, *ins = src_sig.p, *insk = src_skew.p;
        const int NX = nx, vert = vertical, prop = propagation;
        map(N, [=] __device__(size_t q) {
            const int i = (int)(q / NX), j = (int)(q - (size_t)i * NX);
            const T s = vert ? ins[i] : ins[j];
            const T sk = vert ? insk[i] : insk[j];
            if (prop == 0) { S[q] = cmk<C>(s * ng[q], 0); Sk[q] = cmk<C>(-1 * sk * sg[q], 0); }
            else if (prop == 1) { S[q] = cmk<C>(s * ng[q], 0); Sk[q] = cmk<C>(1 * sk * sg[q], 0); }
            else { S[q] = cmk<C>(2 * s * ng[q], 0); Sk[q] = cmk<C>(0, 0); }
        });
        fwd(S);
        fwd(Sk);
        map(N, [=] __device__(size_t q) { S[q] = cadd(S[q], Sk[q]); });
    }

    // ------------------------------------------------------------ RHS --
    // in/out: [eta^][u^][v^], N complex values each, [ny][nx] row-major
    static constexpr long TIME_EVERY = 4;
    int rhs(double t, const C* in, C* out) {
        const bool timed = st.rhs_evals % TIME_EVERY == 0;
        if (timed) {
            if (nev == NEV) { int rc = collect_timing(); if (rc) return rc; }
            HW_TRY(hipEventRecord(ev_a[nev], stream));
        }
        const size_t n = N;
        const int NX = nx;
        const T G = (T)g, nn = (T)n;
        const T *kx = kxv.p, *ky = kyv.p, *fb = fbdy7.p, *w = wall.p;
        const C* Zeta = in;
        const C* Zu = in + n;
        const C* Zv = in + 2 * n;
        C *eta = W[0], *u = W[1], *v = W[2], *dampeta = W[3], *dampu = W[4], *dampv = W[5];
        C *ikxE_adj = W[6], *ikyE_adj = W[7], *duH2_adj = W[8], *dvH2_adj = W[9], *delphiH1 = W[10];
        C *duH1 = W[11], *dvH1 = W[12], *etahat = W[13], *uhat = W[14], *vhat = W[15];
        C *Lphi = W[16], *M0u = W[17], *M0v = W[18];
        if (flat) {
            // fun_delphiH1hat (2d/fun_odesolver.c:2100-2239)
            const T* c2 = c2d.p;
            if (!wall_on) {
                map(n, [=] __device__(size_t q) {
                    const int i = (int)(q / NX), j = (int)(q - (size_t)i * NX);
                    const T cc = c2[q] * c2[q] / G;
                    const T dur = cc * Zu[q].x, dui = cc * Zu[q].y, dvr = cc * Zv[q].x, dvi = cc * Zv[q].y;
                    delphiH1[q] = cmk<C>(kx[j] * dui + ky[i] * dvi, -kx[j] * dur - ky[i] * dvr);
                });
            } else if (refl_one) {
                C *Cx = W[19], *Cy = W[20], *Fx = W[21], *Fy = W[22], *dv = W[23], *a = W[24], *b = W[25];
                const T *gx = gwx.p, *gy = gwy.p;
                map(n, [=] __device__(size_t q) {
                    const T cc = c2[q] * c2[q] / G;
                    Cx[q] = cmk<C>(cc * Zu[q].x, cc * Zu[q].y);
                    Cy[q] = cmk<C>(cc * Zv[q].x, cc * Zv[q].y);
                });
                // ifftw_2d_c2r projects its input in place
                proj(Cx);
                proj(Cy);
                copy(Fx, Cx);
                copy(Fy, Cy);
                bwd(Fx);
                bwd(Fy);
                map(n, [=] __device__(size_t q) {
                    const int i = (int)(q / NX), j = (int)(q - (size_t)i * NX);
                    dv[q] = cmk<C>(-(-kx[j] * Cx[q].y - ky[i] * Cy[q].y), -(kx[j] * Cx[q].x + ky[i] * Cy[q].x));
                });
                proj(dv);
                bwd(dv);
                map(n, [=] __device__(size_t q) {
                    a[q] = cmk<C>(dv[q].x / nn * w[q], 0);
                    b[q] = cmk<C>(Fx[q].x / nn * gx[q] + Fy[q].x / nn * gy[q], 0);
                });
                fwd(a);
                fwd(b);
                map(n, [=] __device__(size_t q) { delphiH1[q] = csub(a[q], b[q]); });
            } else {
                C *Cx = W[19], *Cy = W[20], *a = W[24], *b = W[25];
                map(n, [=] __device__(size_t q) {
                    Cx[q] = cmk<C>(c2[q] * Zu[q].x, c2[q] * Zu[q].y);
                    Cy[q] = cmk<C>(c2[q] * Zv[q].x, c2[q] * Zv[q].y);
                });
                proj(Cx);
                proj(Cy);
                bwd(Cx);
                bwd(Cy);
                map(n, [=] __device__(size_t q) {
                    a[q] = cmk<C>(Cx[q].x / nn * w[q], 0);
                    b[q] = cmk<C>(Cy[q].x / nn * w[q], 0);
                });
                fwd(a);
                fwd(b);
                map(n, [=] __device__(size_t q) {
                    const int i = (int)(q / NX), j = (int)(q - (size_t)i * NX);
                    const T dur = c2[q] / G * a[q].x, dui = c2[q] / G * a[q].y;
                    const T dvr = c2[q] / G * b[q].x, dvi = c2[q] / G * b[q].y;
                    delphiH1[q] = cmk<C>(kx[j] * dui + ky[i] * dvi, -kx[j] * dur - ky[i] * dvr);
                });
            }
            // fun_termshat_AB2_2DF_uv head: M0 by C2d^2/g (-i k); the AB2
            // body's spectra are the input's
            map(n, [=] __device__(size_t q) {
                const int i = (int)(q / NX), j = (int)(q - (size_t)i * NX);
                const T cc = c2[q] * c2[q] / G;
                M0u[q] = cmk<C>(cc * kx[j] * Zu[q].y, -cc * kx[j] * Zu[q].x);
                M0v[q] = cmk<C>(cc * ky[i] * Zv[q].y, -cc * ky[i] * Zv[q].x);
                Lphi[q] = cadd(M0u[q], M0v[q]);
                etahat[q] = Zeta[q];
                uhat[q] = Zu[q];
                vhat[q] = Zv[q];
            });
        } else {
            // fun_termshat_AB1_2DB_uv (2d/fun_odesolver.c:2874-2960)
            copy(eta, Zeta);
            copy(uhat, Zu);
            copy(vhat, Zv);
            proj(uhat);
            proj(vhat);
            proj(eta);
            copy(u, uhat);
            copy(v, vhat);
            bwd(eta);
            bwd(u);
            bwd(v);
            C *Cu = W[19], *Cv = W[20];
            hss(u, uhat, csq_m.p, csq_c.p, csq_p.p, Cu, (double)n);
            hss(v, vhat, csq_m.p, csq_c.p, csq_p.p, Cv, (double)n);
            map(n, [=] __device__(size_t q) {
                dampeta[q] = cmk<C>(fb[q] * eta[q].x / nn, 0);
                dampu[q] = cmk<C>(fb[q] * u[q].x / nn, 0);
                dampv[q] = cmk<C>(fb[q] * v[q].x / nn, 0);
            });
            map(n, [=] __device__(size_t q) {
                const int i = (int)(q / NX), j = (int)(q - (size_t)i * NX);
                duH1[q] = cmk<C>(Cu[q].y / G * kx[j], -Cu[q].x / G * kx[j]);
                dvH1[q] = cmk<C>(Cv[q].y / G * ky[i], -Cv[q].x / G * ky[i]);
            });
            fwd(dampeta);
            fwd(dampu);
            fwd(dampv);
            // fun_termshat_AB2_2DB_uv head
            map(n, [=] __device__(size_t q) {
                etahat[q] = Zeta[q];
                uhat[q] = Zu[q];
                vhat[q] = Zv[q];
                M0u[q] = duH1[q];
                M0v[q] = dvH1[q];
                Lphi[q] = cadd(duH1[q], dvH1[q]);
            });
        }
        C* Shat = W[26];
        const T* al = aal.p;
        if (hs1) {
            // rhs_AB1_2DB_uv (2d/fun_odesolver.c:4350-4402) / rhs_AB1_2DF_uv (:4404-4456)
            if (flat) {
                copy(eta, Zeta);
                copy(u, Zu);
                copy(v, Zv);
                proj(eta);
                proj(u);
                proj(v);
                bwd(eta);
                bwd(u);
                bwd(v);
                map(n, [=] __device__(size_t q) {
                    dampeta[q] = cmk<C>(fb[q] * eta[q].x / nn, 0);
                    dampu[q] = cmk<C>(fb[q] * u[q].x / nn, 0);
                    dampv[q] = cmk<C>(fb[q] * v[q].x / nn, 0);
                });
                fwd(dampeta);
                fwd(dampu);
                fwd(dampv);
            }
            source_hat(t, Shat);
            const bool fl = flat;
            C *oe = out, *ou = out + n, *ov = out + 2 * n;
            map(n, [=] __device__(size_t q) {
                const int i = (int)(q / NX), j = (int)(q - (size_t)i * NX);
                const T a = al[q];
                const C d = fl ? delphiH1[q] : cadd(duH1[q], dvH1[q]);
                oe[q] = cmk<C>((Shat[q].x + d.x - dampeta[q].x) * a, (Shat[q].y + d.y - dampeta[q].y) * a);
                ou[q] = cmk<C>((G * kx[j] * Zeta[q].y - dampu[q].x) * a, (-G * kx[j] * Zeta[q].x - dampu[q].y) * a);
                ov[q] = cmk<C>((G * ky[i] * Zeta[q].y - dampv[q].x) * a, (-G * ky[i] * Zeta[q].x - dampv[q].y) * a);
            });
            return finish_rhs(timed);
        }
        // ---- shared AB2 body (2DF: 2d/fun_odesolver.c:3182-3303, 2DB: 3372-3487)
        proj(etahat);
        proj(uhat);
        proj(vhat);
        proj(Lphi);
        proj(M0u);
        proj(M0v);
        copy(eta, etahat);
        copy(u, uhat);
        copy(v, vhat);
        bwd(eta);
        bwd(u);
        bwd(v);
        bwd(Lphi);
        bwd(M0u);
        bwd(M0v);
        if (flat) {
            map(n, [=] __device__(size_t q) {
                dampeta[q] = cmk<C>(fb[q] * eta[q].x / nn, 0);
                dampu[q] = cmk<C>(fb[q] * u[q].x / nn, 0);
                dampv[q] = cmk<C>(fb[q] * v[q].x / nn, 0);
            });
            fwd(dampeta);
            fwd(dampu);
            fwd(dampv);
        }
        const T N2 = nn * nn;
        C *eM0u = W[27], *eu = W[28], *eM0v = W[29], *ev = W[30], *dE = W[31];
        C *eM0uh = W[32], *euh = W[33], *eM0vh = W[34], *evh = W[35], *dEh = W[36];
        map(n, [=] __device__(size_t q) {
            const T ww = w[q];
            eM0u[q] = cmk<C>(eta[q].x * M0u[q].x / N2 * ww, 0);
            eu[q] = cmk<C>(eta[q].x * u[q].x / N2 * ww, 0);
            eM0v[q] = cmk<C>(eta[q].x * M0v[q].x / N2 * ww, 0);
            ev[q] = cmk<C>(eta[q].x * v[q].x / N2 * ww, 0);
            dE[q] = cmk<C>((T)0.5 * ((u[q].x * u[q].x + v[q].x * v[q].x) - Lphi[q].x * Lphi[q].x) / N2 * ww, 0);
            eM0uh[q] = eM0u[q];
            euh[q] = eu[q];
            eM0vh[q] = eM0v[q];
            evh[q] = ev[q];
            dEh[q] = dE[q];
        });
        fwd(eM0uh);
        fwd(euh);
        fwd(eM0vh);
        fwd(evh);
        fwd(dEh);
        C *LeM0u = W[37], *LeM0v = W[38];
        if (flat) {
            const T* l = l2d.p;
            map(n, [=] __device__(size_t q) {
                LeM0u[q] = cscale(eM0uh[q], l[q]);
                LeM0v[q] = cscale(eM0vh[q], l[q]);
            });
        } else {
            map(n, [=] __device__(size_t q) {
                eM0u[q].x *= nn;
                eM0v[q].x *= nn;
            });
            hss(eM0u, eM0uh, om2_m.p, om2_c.p, om2_p.p, LeM0u, (double)n);
            hss(eM0v, eM0vh, om2_m.p, om2_c.p, om2_p.p, LeM0v, (double)n);
        }
        C *ikxE = W[39], *ikyE = W[40], *duH2 = W[41], *dvH2 = W[42];
        {
            const bool fl = flat;
            map(n, [=] __device__(size_t q) {
                const int i = (int)(q / NX), j = (int)(q - (size_t)i * NX);
                ikxE[q] = cmk<C>(-kx[j] * dEh[q].y, kx[j] * dEh[q].x);
                ikyE[q] = cmk<C>(-ky[i] * dEh[q].y, ky[i] * dEh[q].x);
                if (fl) {
                    duH2[q] = cmk<C>(-LeM0u[q].x + kx[j] * euh[q].y, -LeM0u[q].y - kx[j] * euh[q].x);
                    dvH2[q] = cmk<C>(-LeM0v[q].x + ky[i] * evh[q].y, -LeM0v[q].y - ky[i] * evh[q].x);
                } else {
                    duH2[q] = cmk<C>(kx[j] * euh[q].y - LeM0u[q].x / G, -kx[j] * euh[q].x - LeM0u[q].y / G);
                    dvH2[q] = cmk<C>(ky[i] * evh[q].y - LeM0v[q].x / G, -ky[i] * evh[q].x - LeM0v[q].y / G);
                }
            });
        }
        if (hs3) hs3_terms(in, eta, M0u, M0v, Lphi, LeM0u, LeM0v, duH2, dvH2, ikxE, ikyE);
        adjust(duH2_adj, duH2);
        adjust(dvH2_adj, dvH2);
        adjust(ikxE_adj, ikxE);
        adjust(ikyE_adj, ikyE);
        // ---- source + epilogue (2d/fun_odesolver.c:4479-4529 / 4564-4613)
        source_hat(t, Shat);
        C* dph = W[43];
        {
            const bool fl = flat;
            map(n, [=] __device__(size_t q) {
                const C d = fl ? delphiH1[q] : cadd(duH1[q], dvH1[q]);
                dph[q] = cadd(cadd(d, duH2_adj[q]), dvH2_adj[q]);
            });
        }
        // 2d/fun_odesolver.c:4580-4587: breaking only when tprev < t; Sb^ stays
        // as the last call that ran it left it
        const bool brk_now = brk_on && tprev < t;
        if (brk_on) last_gated = !brk_now;
        if (brk_now) {
            int rc = breaking(t, Zeta, eta, u, v, dph);
            if (rc) return rc;
        }
        const bool sb = brk_on;
        const C* sx = sb ? Sbx.p : nullptr;
        const C* sy = sb ? Sby.p : nullptr;
        C *oe = out, *ou = out + n, *ov = out + 2 * n;
        map(n, [=] __device__(size_t q) {
            const int i = (int)(q / NX), j = (int)(q - (size_t)i * NX);
            const T a = al[q];
            const C bx = sb ? sx[q] : cmk<C>(0, 0), by = sb ? sy[q] : cmk<C>(0, 0);
            oe[q] = cmk<C>((Shat[q].x + dph[q].x - dampeta[q].x) * a, (Shat[q].y + dph[q].y - dampeta[q].y) * a);
            ou[q] = cmk<C>((G * kx[j] * Zeta[q].y + bx.x - ikxE_adj[q].x - dampu[q].x) * a,
                           (-G * kx[j] * Zeta[q].x + bx.y - ikxE_adj[q].y - dampu[q].y) * a);
            ov[q] = cmk<C>((G * ky[i] * Zeta[q].y + by.x - ikyE_adj[q].x - dampv[q].x) * a,
                           (-G * ky[i] * Zeta[q].x + by.y - ikyE_adj[q].y - dampv[q].y) * a);
        });
        return finish_rhs(timed);
    }

    int finish_rhs(bool timed) {
        HW_LAUNCH_CHECK();
        if (timed) {
            HW_TRY(hipEventRecord(ev_b[nev], stream));
            nev++;
            st.rhs_timed++;
        }
        st.rhs_evals++;
        return HW_OK;
    }

    // The third-order terms of fun_termshat_AB3_2DF_uv (2d/fun_odesolver.c:
    // 3512-3897) / fun_termshat_AB3_2DB_uv (:3899-4349): deluH3, delvH3 and
    // i k deletaH3, added to the HS2 adjust inputs (fun_adjust is linear).
    void hs3_terms(const C* in, const C* eta, const C* M0u, const C* M0v, const C* Lphi, const C* LeM0u,
                   const C* LeM0v, C* duH2, C* dvH2, C* ikxE, C* ikyE) {
        const size_t n = N;
        const int NX = nx;
        const T G = (T)g, nn = (T)n;
        const T *kx = kxv.p, *ky = kyv.p, *w = wall.p;
        const C* Zu = in + n;
        const C* Zv = in + 2 * n;
        C *LM0u = W[44], *LM0v = W[45], *LLphi = W[46], *dgp = W[47], *dxu = W[48], *dyv = W[49];
        const T sc = flat ? (T)1 : (T)1 / G;   // LetaM0*_hat (2DB: HSS / grav)
        map(n, [=] __device__(size_t q) {
            const int i = (int)(q / NX), j = (int)(q - (size_t)i * NX);
            LM0u[q] = cscale(LeM0u[q], sc);
            LM0v[q] = cscale(LeM0v[q], sc);
            LLphi[q] = cadd(LM0u[q], LM0v[q]);
            dgp[q] = cmk<C>(-kx[j] * Zu[q].y - ky[i] * Zv[q].y, kx[j] * Zu[q].x + ky[i] * Zv[q].x);
            dxu[q] = cmk<C>(-kx[j] * Zu[q].y, kx[j] * Zu[q].x);
            dyv[q] = cmk<C>(-ky[i] * Zv[q].y, ky[i] * Zv[q].x);
        });
        C* ph[6] = {LM0u, LM0v, LLphi, dgp, dxu, dyv};
        for (int f = 0; f < 6; ++f) {
            proj(ph[f]);
            bwd(ph[f]);
        }
        // physical products (real parts; the transforms above are unnormalised)
        C *e2M0u = W[19], *e2M0v = W[20], *eLu = W[21], *eLv = W[22], *e2dxu = W[23], *e2dyv = W[24], *dE3 = W[25];
        C *h0 = W[27], *h1 = W[28], *h2 = W[29], *h3 = W[30], *h4 = W[31], *h5 = W[32], *h6 = W[33];
        const T N3 = nn * nn * nn;
        map(n, [=] __device__(size_t q) {
            const T ww = w[q], e = eta[q].x;
            e2M0u[q] = cmk<C>(e * e * M0u[q].x / N3 * ww, 0);
            e2M0v[q] = cmk<C>(e * e * M0v[q].x / N3 * ww, 0);
            eLu[q] = cmk<C>(e / nn * LM0u[q].x / nn, 0);
            eLv[q] = cmk<C>(e / nn * LM0v[q].x / nn, 0);
            e2dxu[q] = cmk<C>(e * e * dxu[q].x / N3 * ww, 0);
            e2dyv[q] = cmk<C>(e * e * dyv[q].x / N3 * ww, 0);
            dE3[q] = cmk<C>(Lphi[q].x / nn * e / nn * dgp[q].x / nn + Lphi[q].x / nn * LLphi[q].x / nn, 0);
            h0[q] = e2M0u[q]; h1[q] = e2M0v[q]; h2[q] = eLu[q]; h3[q] = eLv[q];
            h4[q] = e2dxu[q]; h5[q] = e2dyv[q]; h6[q] = dE3[q];
        });
        C* hat[7] = {h0, h1, h2, h3, h4, h5, h6};
        for (int f = 0; f < 7; ++f) fwd(hat[f]);
        C *LLu = W[34], *LLv = W[35], *L2x = W[36], *L2y = W[43];
        if (flat) {
            const T* l = l2d.p;
            map(n, [=] __device__(size_t q) {
                LLu[q] = cscale(h2[q], l[q]);
                LLv[q] = cscale(h3[q], l[q]);
                L2x[q] = cscale(h4[q], l[q]);
                L2y[q] = cscale(h5[q], l[q]);
            });
        } else {
            map(n, [=] __device__(size_t q) {
                eLu[q].x *= nn;
                eLv[q].x *= nn;
                e2dxu[q].x *= nn;
                e2dyv[q].x *= nn;
            });
            hss(eLu, h2, om2_m.p, om2_c.p, om2_p.p, LLu, (double)n);
            hss(eLv, h3, om2_m.p, om2_c.p, om2_p.p, LLv, (double)n);
            hss(e2dxu, h4, om2_m.p, om2_c.p, om2_p.p, L2x, (double)n);
            hss(e2dyv, h5, om2_m.p, om2_c.p, om2_p.p, L2y, (double)n);
            map(n, [=] __device__(size_t q) {
                LLu[q] = cscale(LLu[q], (T)1 / G);
                LLv[q] = cscale(LLv[q], (T)1 / G);
                L2x[q] = cscale(L2x[q], (T)1 / G);
                L2y[q] = cscale(L2y[q], (T)1 / G);
            });
        }
        map(n, [=] __device__(size_t q) {
            const int i = (int)(q / NX), j = (int)(q - (size_t)i * NX);
            const T kx2 = kx[j] * kx[j], ky2 = ky[i] * ky[i];
            duH2[q].x += LLu[q].x + (T)0.5 * L2x[q].x - (T)0.5 * (kx2 * h0[q].x);
            duH2[q].y += LLu[q].y + (T)0.5 * L2x[q].y - (T)0.5 * (kx2 * h0[q].y);
            dvH2[q].x += LLv[q].x + (T)0.5 * L2y[q].x - (T)0.5 * (ky2 * h1[q].x);
            dvH2[q].y += LLv[q].y + (T)0.5 * L2y[q].y - (T)0.5 * (ky2 * h1[q].y);
            ikxE[q].x += -kx[j] * h6[q].y;
            ikxE[q].y += kx[j] * h6[q].x;
            ikyE[q].x += -ky[i] * h6[q].y;
            ikyE[q].y += ky[i] * h6[q].x;
        });
    }

    // breaking_process (2d/breaking.c:505-733) and fun_sbreaking
    // (2d/fun_odesolver.c:985-1019) at time t; dph (delphiH) is projected in
    // place, as fun_sbreaking does
    int breaking(double t, const C* Zeta, const C* eta, const C* u, const C* v, C* dph) {
        const size_t n = N;
        const int NX = nx, NY = ny;
        const T nn = (T)n;
        T *e = ep.p, *uu = up.p, *vv = vp.p, *H = hil.p;
        map(n, [=] __device__(size_t q) {
            e[q] = eta[q].x / nn;
            uu[q] = u[q].x / nn;
            vv[q] = v[q].x / nn;
        });
        int rc;
        if ((rc = cm.detect_begin(stream, e))) return rc;
        // Hilbert transform of eta: F^-1(P(-i sign(kx + ky) eta^)) (2d/breaking.c:256-261)
        C* Hc = W[54];
        const T *kx = kxv.p, *ky = kyv.p;
        map(n, [=] __device__(size_t q) {
            const int i = (int)(q / NX), j = (int)(q - (size_t)i * NX);
            const T sg = (kx[j] + ky[i]) < 0 ? (T)-1 : (T)1;
            Hc[q] = cmk<C>(Zeta[q].y * sg, -Zeta[q].x * sg);
        });
        proj(Hc);
        bwd(Hc);
        map(n, [=] __device__(size_t q) { H[q] = Hc[q].x / nn; });
        HW_LAUNCH_CHECK();
        if ((rc = cm.detect_end(stream, e, H, uu, vv))) return rc;
        bool have_b = false;
        if ((rc = cm.run(t, stream, e, uu, vv, &have_b))) return rc;
        // fun_sbreaking
        proj(dph);
        C* sxc = Sbx.p;
        C* syc = Sby.p;
        if (!have_b) {
            // B == 0: the flux and its gradient vanish
            map(n, [=] __device__(size_t q) { sxc[q] = cmk<C>(0, 0); syc[q] = cmk<C>(0, 0); });
            HW_LAUNCH_CHECK();
            return HW_OK;
        }
        C* d = W[55];
        copy(d, dph);
        bwd(d);
        const T db2 = (T)(delb * delb);
        const T *B = cm.Bf.p, *D = cm.depth.p;
        T* F = flux.p;
        map(n, [=] __device__(size_t q) {
            const T dp = d[q].x / nn;
            F[q] = -B[q] * db2 * (e[q] + D[q]) * dp * dp;
        });
        k_brk_sb<T><<<dim3((nx + 255) / 256, ny), 256, 0, stream>>>(F, e, D, cm.dxg, cm.dyg, nx, ny, sbxp.p, sbyp.p);
        const T *px = sbxp.p, *py = sbyp.p;
        map(n, [=] __device__(size_t q) { sxc[q] = cmk<C>(px[q], 0); syc[q] = cmk<C>(py[q], 0); });
        fwd(sxc);
        fwd(syc);
        HW_LAUNCH_CHECK();
        return HW_OK;
    }

    // ------------------------------------------------ boundary layout --
    // 6 planes [Re eta^][Im eta^][Re u^][Im u^][Re v^][Im v^] <-> 3 complex arrays
    template <typename CO> void planes_to(const double* d, CO* o) {
        const size_t n = N;
        map(3 * n, [=] __device__(size_t q) {
            const size_t f = q / n, r = q - f * n;
            CO z;
            z.x = (decltype(z.x))d[2 * f * n + r];
            z.y = (decltype(z.y))d[(2 * f + 1) * n + r];
            o[q] = z;
        });
    }
    template <typename CI> void to_planes(const CI* a, double* d) {
        const size_t n = N;
        map(3 * n, [=] __device__(size_t q) {
            const size_t f = q / n, r = q - f * n;
            d[2 * f * n + r] = (double)a[q].x;
            d[(2 * f + 1) * n + r] = (double)a[q].y;
        });
    }

    int rhs_full_dev(double t, const double* d_in, double* d_out) override {
        planes_to<C>(d_in, ytmp.p);
        int rc = rhs(t, ytmp.p, kbuf.p);
        if (rc) return rc;
        to_planes<C>(kbuf.p, d_out);
        HW_LAUNCH_CHECK();
        HW_TRY(hipStreamSynchronize(stream));
        return collect_timing();
    }
    int set_state_dev(const double* d_full) override {
        planes_to<double2>(d_full, y.p);
        HW_LAUNCH_CHECK();
        HW_TRY(hipStreamSynchronize(stream));
        return HW_OK;
    }
    int get_state_dev(double* d_full) override {
        to_planes<double2>(y.p, d_full);
        HW_LAUNCH_CHECK();
        HW_TRY(hipStreamSynchronize(stream));
        return HW_OK;
    }

    // ---------------------------------------------------------- RKF45 --
    C* K(int i) { return kbuf.p + (size_t)i * 3 * N; }
    int stage(double h, std::initializer_list<int> ks, std::initializer_list<double> bs) {
        StageCoef sc{};
        sc.nk = 0;
        auto bi = bs.begin();
        for (int kidx : ks) {
            sc.k[sc.nk] = K(kidx);
            sc.b[sc.nk] = *bi++;
            sc.nk++;
        }
        k_rk_stage<T><<<grid1(3 * N), 256, 0, stream>>>(y.p, sc, h, ytmp.p, 3 * N);
        HW_LAUNCH_CHECK();
        return HW_OK;
    }
    // one attempted rkf45 step (see Solver2DT::attempt_issue); the error norm
    // covers all 6 nx ny values of the full state, as GSL's does
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
        k_rk_final<T><<<dim3(1024), 256, 0, stream>>>(y.p, K(0), K(2), K(3), K(4), K(5), h0, eps_abs, eps_rel, ynew.p,
                                                        rmax.p, 3 * N, brk_on ? ytmp.p : nullptr);
        HW_LAUNCH_CHECK();
        if (brk_on) {
            // rkf45's dydt_out = f(t + h, y_new) with its breaking side effects
            if ((rc = rhs(t0 + h0, ytmp.p, K(1)))) return rc;
        }
        HW_TRY(hipMemcpyAsync(h_rmax, rmax.p, sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
        return HW_OK;
    }
    int attempt_result(double* rm) override {
        HW_TRY(hipStreamSynchronize(stream));
        std::memcpy(rm, h_rmax, sizeof(double));
        if (*rm < 2.2250738585072014e-308) *rm = 2.2250738585072014e-308;
        return HW_OK;
    }
    int accept_step(double t) override {
        std::swap(y.p, ynew.p);
        st.steps_accepted++;
        tprev = t;
        return collect_timing();
    }
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

    // ------------------------------------------------------- outputs --
    // physical eta, u, v of the state: real part of the projected inverse / N
    // (the output IFFTs of ode_solver_uv, 2d/fun_odesolver.c:2663-2666)
    int fields_dev(double* eta, double* u, double* v) override {
        double* outs[3] = {eta, u, v};
        const size_t n = N;
        for (int f = 0; f < 3; ++f) {
            if (!outs[f]) continue;
            C* a = W[0];
            const double2* s = y.p + (size_t)f * n;
            map(n, [=] __device__(size_t q) { a[q] = cmk<C>((T)s[q].x, (T)s[q].y); });
            proj(a);
            bwd(a);
            double* o = outs[f];
            const double nd = (double)n;
            map(n, [=] __device__(size_t q) { o[q] = (double)a[q].x / nd; });
            HW_LAUNCH_CHECK();
        }
        HW_TRY(hipStreamSynchronize(stream));
        return HW_OK;
    }

    int breaking_state(double* Bout, int* fl) override {
        if (fl) *fl = cm.flagbr;
        if (Bout) {
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

    // graph mode and members are features of the fused path
    int graph_begin(const double*, double, const double*) override {
        set_error("device step control (HW_STEP_GRAPH) is not available on the full-grid path");
        return HW_ERR_ARG;
    }
    int graph_launch() override { return HW_ERR_ARG; }
    int graph_poll(bool*) override { return HW_ERR_ARG; }
    int graph_finish(double*, double*) override { return HW_ERR_ARG; }
    int set_member_dev(const double*, int) override { set_error("the full-grid path holds one state"); return HW_ERR_ARG; }
    int get_member_dev(double*, int) override { set_error("the full-grid path holds one state"); return HW_ERR_ARG; }
};

}  // namespace hw
