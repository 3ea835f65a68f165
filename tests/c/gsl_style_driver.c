/* A C host that uses libhawassi_hip.so the way the reference's ode_solver_uv
 * uses its right-hand side (2d/fun_odesolver.c:2565-2647, 2d/main.c:204-222):
 * the RHS is registered as the function of a gsl_odeiv2_system-shaped
 * struct, {hw_rhs, NULL, 6*nx*ny, ctx}, and a host-side rkf45 step with
 * control_y(eps_abs, eps_rel) and evolve_apply advances the full 6-plane
 * host state, calling only that function pointer.  GSL is not part of this
 * image, so the stepper below restates GSL 2.x odeiv2's rkf45 / control_y /
 * evolve_apply (Fehlberg 4(5), the 5th-order solution propagated; ord = 5;
 * decrease by max(0.9 r^-1/5, 0.2) when r > 1.1, increase by
 * min(max(0.9 r^-1/6, 1), 5) when r < 0.5).
 *
 *   gsl_style_driver MODEL_FILE OUT_FILE
 *
 * MODEL_FILE (little-endian, written by the test that runs this program):
 *   int32  nx, ny, flat, model, precision
 *   double grav, cpeak, eps_abs, eps_rel, t0, t1, h0
 *   double kx[nx], ky[ny], aal, fbdy, wallchar, chi [ny*nx]
 *   flat:  double c2d, l2d [ny*nx]
 *   else:  double gam_min, gam_plus, csq_min, csq_plus, om2_min, om2_plus [ny*nx]
 *   double y0[6*nx*ny]
 * No influx source and no breaking (an autonomous basin).
 * OUT_FILE: double y[6*nx*ny] at t1, then long accepted, rejected, rhs calls.
 * Compiles as C99 against include/hawassi_hip.h only. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hawassi_hip.h"

/* the shape of gsl_odeiv2_system */
typedef struct {
    int (*function)(double t, const double y[], double dydt[], void *params);
    int (*jacobian)(double t, const double y[], double *dfdy, double dfdt[], void *params);
    size_t dimension;
    void *params;
} odeiv2_system;

typedef struct {
    size_t n;
    double *k[6], *ytmp, *y0, *yerr, *dydt_in, *dydt_out;
    long accepted, rejected, calls;
} stepper;

static int call(stepper *s, const odeiv2_system *sys, double t, const double *y, double *dydt)
{
    s->calls++;
    return sys->function(t, y, dydt, sys->params);
}

static int rkf45_apply(stepper *s, const odeiv2_system *sys, double t, double h, double *y)
{
    static const double ah[5] = {1.0 / 4.0, 3.0 / 8.0, 12.0 / 13.0, 1.0, 1.0 / 2.0};
    static const double b3[2] = {3.0 / 32.0, 9.0 / 32.0};
    static const double b4[3] = {1932.0 / 2197.0, -7200.0 / 2197.0, 7296.0 / 2197.0};
    static const double b5[4] = {8341.0 / 4104.0, -32832.0 / 4104.0, 29440.0 / 4104.0, -845.0 / 4104.0};
    static const double b6[5] = {-6080.0 / 20520.0, 41040.0 / 20520.0, -28352.0 / 20520.0, 9295.0 / 20520.0,
                                 -5643.0 / 20520.0};
    static const double c1 = 902880.0 / 7618050.0, c3 = 3953664.0 / 7618050.0, c4 = 3855735.0 / 7618050.0,
                        c5 = -1371249.0 / 7618050.0, c6 = 277020.0 / 7618050.0;
    static const double ec[7] = {0.0, 1.0 / 360.0, 0.0, -128.0 / 4275.0, -2197.0 / 75240.0, 1.0 / 50.0, 2.0 / 55.0};
    size_t i, n = s->n;
    double **k = s->k, *yt = s->ytmp;
    int rc;
    memcpy(k[0], s->dydt_in, sizeof(double) * n);
    for (i = 0; i < n; i++) yt[i] = y[i] + ah[0] * h * k[0][i];
    if ((rc = call(s, sys, t + ah[0] * h, yt, k[1]))) return rc;
    for (i = 0; i < n; i++) yt[i] = y[i] + h * (b3[0] * k[0][i] + b3[1] * k[1][i]);
    if ((rc = call(s, sys, t + ah[1] * h, yt, k[2]))) return rc;
    for (i = 0; i < n; i++) yt[i] = y[i] + h * (b4[0] * k[0][i] + b4[1] * k[1][i] + b4[2] * k[2][i]);
    if ((rc = call(s, sys, t + ah[2] * h, yt, k[3]))) return rc;
    for (i = 0; i < n; i++)
        yt[i] = y[i] + h * (b5[0] * k[0][i] + b5[1] * k[1][i] + b5[2] * k[2][i] + b5[3] * k[3][i]);
    if ((rc = call(s, sys, t + ah[3] * h, yt, k[4]))) return rc;
    for (i = 0; i < n; i++)
        yt[i] = y[i] + h * (b6[0] * k[0][i] + b6[1] * k[1][i] + b6[2] * k[2][i] + b6[3] * k[3][i] + b6[4] * k[4][i]);
    if ((rc = call(s, sys, t + ah[4] * h, yt, k[5]))) return rc;
    for (i = 0; i < n; i++)
        y[i] += h * (c1 * k[0][i] + c3 * k[2][i] + c4 * k[3][i] + c5 * k[4][i] + c6 * k[5][i]);
    /* GSL's rkf45 evaluates dydt_out at the new point when evolve_apply
       passes one (it always does) */
    if ((rc = call(s, sys, t + h, y, s->dydt_out))) return rc;
    for (i = 0; i < n; i++)
        s->yerr[i] = h * (ec[1] * k[0][i] + ec[3] * k[2][i] + ec[4] * k[3][i] + ec[5] * k[4][i] + ec[6] * k[5][i]);
    return 0;
}

/* control_y: -1 decrease, 0 keep, +1 increase */
static int hadjust(const stepper *s, const double *y, double eps_abs, double eps_rel, double *h)
{
    const double ord = 5.0;
    double rmax = 2.2250738585072014e-308;
    size_t i;
    for (i = 0; i < s->n; i++) {
        double r = fabs(s->yerr[i]) / fabs(eps_rel * fabs(y[i]) + eps_abs);
        if (r > rmax) rmax = r;
    }
    if (rmax > 1.1) {
        double r = 0.9 / pow(rmax, 1.0 / ord);
        *h *= r < 0.2 ? 0.2 : r;
        return -1;
    }
    if (rmax < 0.5) {
        double r = 0.9 / pow(rmax, 1.0 / (ord + 1.0));
        if (r > 5.0) r = 5.0;
        if (r < 1.0) r = 1.0;
        *h *= r;
        return 1;
    }
    return 0;
}

/* evolve_apply: one accepted step (the last one clipped to t1) */
static int evolve_apply(stepper *s, const odeiv2_system *sys, double *t, double t1, double *h, double *y,
                        double eps_abs, double eps_rel)
{
    const double t0 = *t;
    double h0 = *h, dt = t1 - t0;
    int rc;
    memcpy(s->y0, y, sizeof(double) * s->n);
    if ((rc = call(s, sys, t0, y, s->dydt_in))) return rc;
    for (;;) {
        int final_step = (dt >= 0.0 && h0 > dt) || (dt < 0.0 && h0 < dt);
        double h_old;
        if (final_step) h0 = dt;
        if ((rc = rkf45_apply(s, sys, t0, h0, y))) return rc;
        *t = final_step ? t1 : t0 + h0;
        h_old = h0;
        if (hadjust(s, y, eps_abs, eps_rel, &h0) < 0) {
            if (fabs(h0) < fabs(h_old) && *t + h0 != *t) {
                memcpy(y, s->y0, sizeof(double) * s->n);
                s->rejected++;
                continue;
            }
            *h = h0;
            *t = t0;
            return HW_ERR_STEP;
        }
        s->accepted++;
        *h = h0;
        return 0;
    }
}

static int rd(FILE *f, void *p, size_t size, size_t count)
{
    return fread(p, size, count, f) == count ? 0 : -1;
}

static double *rd_arr(FILE *f, size_t count)
{
    double *a = (double *)malloc(sizeof(double) * count);
    if (!a || rd(f, a, sizeof(double), count)) {
        fprintf(stderr, "gsl_style_driver: short model file\n");
        exit(2);
    }
    return a;
}

int main(int argc, char **argv)
{
    int hdr[6];
    double par[7];
    hw_config_2d cfg;
    odeiv2_system sys;
    stepper s;
    void *ctx = NULL;
    FILE *f;
    size_t nn, dim, i;
    double *y, t, t1, h;
    int rc;

    if (argc != 3) {
        fprintf(stderr, "usage: gsl_style_driver MODEL_FILE OUT_FILE\n");
        return 2;
    }
    f = fopen(argv[1], "rb");
    if (!f || rd(f, hdr, sizeof(int), 6) || rd(f, par, sizeof(double), 7)) {
        fprintf(stderr, "gsl_style_driver: cannot read %s\n", argv[1]);
        return 2;
    }
    memset(&cfg, 0, sizeof(cfg));
    cfg.nx = hdr[0];
    cfg.ny = hdr[1];
    cfg.flat = hdr[2];
    cfg.model = hdr[3];
    cfg.precision = hdr[4];
    cfg.full_grid = hdr[5];   /* 1: the reference's own full-grid values in every column */
    cfg.grav = par[0];
    cfg.cpeak = par[1];
    cfg.eps_abs = par[2];
    cfg.eps_rel = par[3];
    cfg.t0 = par[4];
    nn = (size_t)cfg.nx * (size_t)cfg.ny;
    dim = 6 * nn;
    cfg.kx = rd_arr(f, (size_t)cfg.nx);
    cfg.ky = rd_arr(f, (size_t)cfg.ny);
    cfg.aal = rd_arr(f, nn);
    cfg.fbdy = rd_arr(f, nn);
    cfg.wallchar = rd_arr(f, nn);
    cfg.chi = rd_arr(f, nn);
    if (cfg.flat) {
        cfg.c2d = rd_arr(f, nn);
        cfg.l2d = rd_arr(f, nn);
    } else {
        cfg.gam_min = rd_arr(f, nn);
        cfg.gam_plus = rd_arr(f, nn);
        cfg.csq_min = rd_arr(f, nn);
        cfg.csq_plus = rd_arr(f, nn);
        cfg.om2_min = rd_arr(f, nn);
        cfg.om2_plus = rd_arr(f, nn);
    }
    y = rd_arr(f, dim);
    fclose(f);

    if ((rc = hw_ctx_create_2d(&cfg, &ctx)) != HW_OK) {
        fprintf(stderr, "hw_ctx_create_2d: %s\n", hw_last_error());
        return 1;
    }
    /* ode_solver_uv's system: {rhs, NULL, 6*Nx*Ny, params} */
    sys.function = hw_rhs;
    sys.jacobian = NULL;
    sys.dimension = dim;
    sys.params = ctx;

    memset(&s, 0, sizeof(s));
    s.n = dim;
    for (i = 0; i < 6; i++) s.k[i] = (double *)malloc(sizeof(double) * dim);
    s.ytmp = (double *)malloc(sizeof(double) * dim);
    s.y0 = (double *)malloc(sizeof(double) * dim);
    s.yerr = (double *)malloc(sizeof(double) * dim);
    s.dydt_in = (double *)malloc(sizeof(double) * dim);
    s.dydt_out = (double *)malloc(sizeof(double) * dim);

    /* the driver loop of ode_solver_uv: while (t < t1) evolve_apply */
    t = par[4];
    t1 = par[5];
    h = par[6];
    while (t < t1) {
        if ((rc = evolve_apply(&s, &sys, &t, t1, &h, y, cfg.eps_abs, cfg.eps_rel)) != 0) {
            fprintf(stderr, "evolve_apply failed at t = %g: %d %s\n", t, rc, hw_last_error());
            hw_ctx_destroy(ctx);
            return 1;
        }
    }
    printf("t = %.6f accepted %ld rejected %ld rhs calls %ld\n", t, s.accepted, s.rejected, s.calls);

    f = fopen(argv[2], "wb");
    if (!f || fwrite(y, sizeof(double), dim, f) != dim || fwrite(&s.accepted, sizeof(long), 1, f) != 1 ||
        fwrite(&s.rejected, sizeof(long), 1, f) != 1 || fwrite(&s.calls, sizeof(long), 1, f) != 1) {
        fprintf(stderr, "gsl_style_driver: cannot write %s\n", argv[2]);
        return 2;
    }
    fclose(f);
    hw_ctx_destroy(ctx);
    return 0;
}
