// crest2d.h -- the crest bookkeeping of the 2-D breaking_process
// (2d/breaking.c:505-733, breaking_criterion :216-334), shared by the fused
// half-spectrum path (solver2d.h) and the full-grid path (solver2dg.h).
//
// Every full-grid scan runs on the device (peak search, crest test at the
// peaks, crest re-location and break-node marking: breaking2d.h); the
// flagbr state machine and the new-crest tests (a few scalars per crest) run
// on the host between them.  The caller supplies the physical eta, u, v and,
// between detect_begin and detect_end, the Hilbert transform of eta (its own
// transforms); the machine leaves the eddy-viscosity field B in Bf.
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <vector>

#include "../../include/hawassi_hip.h"
#include "breaking2d.h"

namespace hw {

void set_error(const char* fmt, ...);

// device buffer owned by a context
template <typename T> struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
    int alloc(size_t count) {
        n = count;
        hipError_t e = hipMalloc((void**)&p, sizeof(T) * (count ? count : 1));
        if (e != hipSuccess) {
            set_error("hipMalloc(%zu bytes) failed: %s", sizeof(T) * count, hipGetErrorString(e));
            p = nullptr;
            return HW_ERR_HIP;
        }
        return HW_OK;
    }
    int upload(const std::vector<T>& h, hipStream_t s) {
        int rc = alloc(h.size());
        if (rc) return rc;
        hipError_t e = hipMemcpyAsync(p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s);
        if (e != hipSuccess) {
            set_error("hipMemcpyAsync failed: %s", hipGetErrorString(e));
            return HW_ERR_HIP;
        }
        return HW_OK;
    }
};

#define HW_CM_TRY(expr)                                                              \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) {                                                      \
            ::hw::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),  \
                            __FILE__, __LINE__);                                     \
            return HW_ERR_HIP;                                                       \
        }                                                                            \
    } while (0)

template <typename T> struct CrestMachine {
    struct CrestRec {                      // CBvar without the node arrays
        int xindex, yindex, xNow, yNow, kwadran;
        double Ucrest, theta, time;
    };
    static constexpr int CAPH = 256;       // candidates copied to the host with the counts
    int nx = 0, ny = 0;
    size_t N = 0;
    double KBC = 0, TC = 0, Tstar = 0, Rsearch = 0, MaxEtaInit = 0, Ucur = 0, dxg = 1, dyg = 1;
    int maxpk = 0, maxcr = 0;
    DevBuf<T> pbc, depth, Bf;
    DevBuf<double> xg, yg;
    DevBuf<int> owner;
    // one region for the counts and the candidates, so that one copy brings
    // both back: pkcount[0] peaks, pkcount[1] crest candidates, then cand
    static constexpr size_t CNT_BYTES = 64;
    DevBuf<unsigned char> cc;
    int* pkcount = nullptr;
    BrkCand* cand = nullptr;
    DevBuf<BrkPeak> peaks;
    DevBuf<BrkCrest> crests;
    unsigned char* h_cc = nullptr;         // pinned copy of the first CAPH candidates and the counts
    int* h_cnt = nullptr;
    BrkCand* h_cand = nullptr;
    hipEvent_t ev = nullptr;
    int flagbr = 0;
    std::vector<CrestRec> cbprev;          // CBPrev entries (count = size)
    long calls = 0;
    bool b_zero = true;                    // Bf known to be all zero

    ~CrestMachine() {
        if (h_cc) (void)hipHostFree(h_cc);
        if (ev) (void)hipEventDestroy(ev);
    }

    static dim3 grid1(size_t n) {
        size_t b = (n + 255) / 256;
        if (b > 65535u * 8) b = 65535u * 8;
        return dim3((unsigned)b);
    }

    int init(const hw_config_2d* c, int nx_, int ny_, hipStream_t s) {
        nx = nx_; ny = ny_; N = (size_t)nx * ny;
        if (!c->pbchar || !c->depth || !c->x || !c->y || c->rsearch <= 0) {
            set_error("breaking needs pbchar, depth, x, y and rsearch > 0");
            return HW_ERR_ARG;
        }
        KBC = c->KBC; TC = c->TC; MaxEtaInit = c->max_eta_init;
        Tstar = c->tstar; Rsearch = c->rsearch; Ucur = c->current;
        dxg = c->x[1] - c->x[0]; dyg = c->y[1] - c->y[0];
        std::vector<T> hp(N), hd(N);
        for (size_t i = 0; i < N; ++i) { hp[i] = (T)c->pbchar[i]; hd[i] = (T)c->depth[i]; }
        std::vector<double> xv(c->x, c->x + nx), yv(c->y, c->y + ny);
        int rc;
        if ((rc = pbc.upload(hp, s)) || (rc = depth.upload(hd, s)) || (rc = xg.upload(xv, s)) || (rc = yg.upload(yv, s)))
            return rc;
        maxpk = (int)std::max<size_t>(4096, N / 8);
        maxcr = 65536;
        if ((rc = Bf.alloc(N)) || (rc = owner.alloc(N)) || (rc = cc.alloc(CNT_BYTES + sizeof(BrkCand) * maxpk)) ||
            (rc = peaks.alloc(maxpk)) || (rc = crests.alloc(maxcr)))
            return rc;
        pkcount = reinterpret_cast<int*>(cc.p);
        cand = reinterpret_cast<BrkCand*>(cc.p + CNT_BYTES);
        k_fill<T><<<grid1(N), 256, 0, s>>>(Bf.p, (T)0, N);
        HW_CM_TRY(hipGetLastError());
        HW_CM_TRY(hipHostMalloc((void**)&h_cc, CNT_BYTES + CAPH * sizeof(BrkCand), hipHostMallocDefault));
        h_cnt = reinterpret_cast<int*>(h_cc);
        h_cand = reinterpret_cast<BrkCand*>(h_cc + CNT_BYTES);
        HW_CM_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        return HW_OK;
    }

    // fun_peaks (2d/breaking.c:55-133) on eta * pb.Char; pkcount[0] holds
    // the count for the caller's Hilbert kernels, which may skip their work
    // when it is zero
    int detect_begin(hipStream_t s, const T* eta, bool zeroed = false) {
        // zeroed: an earlier kernel of this call has zeroed the counters
        if (!zeroed) k_fill<int><<<1, 64, 0, s>>>(pkcount, 0, 2);
        k_brk_peaks<T><<<dim3((nx + 255) / 256, ny), 256, 0, s>>>(eta, pbc.p, MaxEtaInit, nx, ny, peaks.p, pkcount,
                                                                   maxpk);
        HW_CM_TRY(hipGetLastError());
        return HW_OK;
    }

    // the crest test at the peaks (breaking_criterion, 2d/breaking.c:253-305)
    // with H the Hilbert transform of eta; counts and the first candidates
    // come back with an asynchronous copy
    int detect_end(hipStream_t s, const T* eta, const T* H, const T* u, const T* v) {
        k_brk_crit<T><<<64, 256, 0, s>>>(peaks.p, pkcount, maxpk, eta, H, u, v, depth.p, Bf.p, dxg, dyg, Ucur, KBC,
                                         nx, ny, cand);
        HW_CM_TRY(hipGetLastError());
        HW_CM_TRY(hipMemcpyAsync(h_cc, cc.p, CNT_BYTES + CAPH * sizeof(BrkCand), hipMemcpyDeviceToHost, s));
        HW_CM_TRY(hipEventRecord(ev, s));
        return HW_OK;
    }

    // The crest lists of this call (2d/breaking.c:526-709) and B.  Waits for
    // detect_end's copies.  *have_b: some node has B > 0 (else Bf is zero).
    int run(double t, hipStream_t s, const T* eta, const T* u, const T* v, bool* have_b) {
        calls++;
        *have_b = false;
        HW_CM_TRY(hipEventSynchronize(ev));
        // detect_end may have run on another stream: order s after it
        HW_CM_TRY(hipStreamWaitEvent(s, ev, 0));
        const int npk = h_cnt[0], ncand = h_cnt[1];
        if (npk > maxpk) { set_error("breaking: %d peaks exceed the buffer of %d", npk, maxpk); return HW_ERR_ARG; }
        // ---- criterion: candidate crests in the reference's peak order
        std::vector<CrestRec> cb;
        std::vector<int> cb_prevnode;
        if (ncand > 0) {
            std::vector<BrkCand> hc(ncand);
            std::copy(h_cand, h_cand + std::min(ncand, CAPH), hc.begin());
            if (ncand > CAPH)
                HW_CM_TRY(hipMemcpy(hc.data() + CAPH, cand + CAPH, sizeof(BrkCand) * (ncand - CAPH),
                                    hipMemcpyDeviceToHost));
            // fun_peaks order: interior x-major, then x=0, x=nx-1, y=0, y=ny-1 edges
            const int NX = nx, NY = ny;
            auto key = [NX, NY](const BrkCand& p) -> long long {
                int cls;
                long long a, b;
                if (p.x >= 1 && p.x <= NX - 2 && p.y >= 1 && p.y <= NY - 2) { cls = 0; a = p.x; b = p.y; }
                else if (p.x == 0) { cls = 1; a = p.y; b = 0; }
                else if (p.x == NX - 1) { cls = 2; a = p.y; b = 0; }
                else if (p.y == 0) { cls = 3; a = p.x; b = 0; }
                else { cls = 4; a = p.x; b = 0; }
                return ((long long)cls << 40) | (a << 20) | b;
            };
            std::vector<int> ord(ncand);
            for (int i = 0; i < ncand; ++i) ord[i] = i;
            std::sort(ord.begin(), ord.end(), [&](int a, int b) { return key(hc[a]) < key(hc[b]); });
            for (int i : ord) {
                CrestRec r;
                r.xindex = r.xNow = hc[i].x;
                r.yindex = r.yNow = hc[i].y;
                r.Ucrest = hc[i].U;
                r.time = t;
                r.theta = hc[i].theta;
                r.kwadran = hc[i].kwadran;
                cb.push_back(r);
                cb_prevnode.push_back(hc[i].prevnode);
            }
        }
        // ---- crest list for this call (2d/breaking.c:526-593)
        const int ncb = (int)cb.size();
        std::vector<CrestRec> cbnew;
        if (ncb > 0 && flagbr == 0) {
            flagbr = 1;
            cbnew = cb;
        } else if (ncb == 0 && flagbr == 2) {
            cbnew = cbprev;
        } else if (ncb > 0 && flagbr == 2) {
            cbnew = cbprev;
            for (int k = 0; k < ncb; ++k) {
                bool isnew = !cb_prevnode[k];   // fun_check_crest: on a previous break node?
                for (const CrestRec& p : cbprev)
                    if (p.xindex == cb[k].xindex && p.yindex == cb[k].yindex) isnew = false;
                if (isnew) cbnew.push_back(cb[k]);
            }
        }
        bool hb = false;
        if (flagbr == 1 || flagbr == 2) {
            const int nnew = (int)cbnew.size();
            std::vector<CrestRec> next;
            std::vector<int> term(nnew, 0);
            if (nnew > 0) {
                if (nnew > maxcr) { set_error("breaking: %d crests exceed %d", nnew, maxcr); return HW_ERR_ARG; }
                std::vector<BrkCrest> hcr(nnew);
                for (int lx = 0; lx < nnew; ++lx) {
                    const CrestRec& r = cbnew[lx];
                    BrkCrest& b = hcr[lx];
                    b.mode = flagbr == 1 ? 1 : 2;
                    b.xNow = flagbr == 1 ? r.xindex : r.xNow;
                    b.yNow = flagbr == 1 ? r.yindex : r.yNow;
                    b.theta = r.theta;
                    b.kwadran = r.kwadran;
                    double UI = r.Ucrest, UF = TC * UI, dt = t - r.time, Us;
                    if (dt >= Tstar) Us = UF;
                    else if (dt >= 0 && dt < Tstar) Us = UI + (dt / Tstar) * (UF - UI);
                    else Us = UI;
                    b.Ustar = Us;
                    b.UF = UF;
                    b.found = 0;
                    b.nb = 0;
                }
                k_fill<T><<<grid1(N), 256, 0, s>>>(Bf.p, (T)0, N);
                k_fill<int><<<grid1(N), 256, 0, s>>>(owner.p, INT_MAX, N);
                HW_CM_TRY(hipMemcpyAsync(crests.p, hcr.data(), sizeof(BrkCrest) * nnew, hipMemcpyHostToDevice, s));
                BrkGeom geo{xg.p, yg.p, Rsearch, nx, ny};
                k_brk_track<T><<<nnew, 256, 0, s>>>(crests.p, nnew, geo, eta, u, v, pbc.p, owner.p);
                k_brk_resolve<T><<<nnew, 256, 0, s>>>(crests.p, nnew, geo, eta, u, v, pbc.p, owner.p, Bf.p);
                HW_CM_TRY(hipGetLastError());
                HW_CM_TRY(hipMemcpyAsync(hcr.data(), crests.p, sizeof(BrkCrest) * nnew, hipMemcpyDeviceToHost, s));
                HW_CM_TRY(hipStreamSynchronize(s));
                for (int lx = 0; lx < nnew; ++lx) {
                    const BrkCrest& b = hcr[lx];
                    if (!b.found) { term[lx] = 1; continue; }
                    CrestRec r = cbnew[lx];
                    r.xNow = b.xNow; r.yNow = b.yNow; r.theta = b.theta; r.kwadran = b.kwadran;
                    if (b.nb == 0) {
                        term[lx] = 1;
                    } else {
                        term[lx] = 0;
                        next.push_back(r);
                        hb = true;
                    }
                }
                for (int lx = 0; lx < nnew; ++lx) {
                    if (term[lx] == 0) { flagbr = 2; break; }
                    flagbr = 0;
                }
            }
            cbprev = next;
        }
        if (!hb) {
            // B == 0 everywhere
            if (!b_zero) k_fill<T><<<grid1(N), 256, 0, s>>>(Bf.p, (T)0, N);
            HW_CM_TRY(hipGetLastError());
            b_zero = true;
            return HW_OK;
        }
        b_zero = false;
        *have_b = true;
        return HW_OK;
    }
};

}  // namespace hw
