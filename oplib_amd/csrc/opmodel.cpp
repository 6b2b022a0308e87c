// opmodel.cpp -- host model layer over the ophip C-ABI (see opmodel.h).
//
// Plain C++; every device operation goes through a function declared in
// ophip.h, in the order the reference's C# model layer issues the matching
// opcuda calls:
//   grid + stencil          SGrid.cs:492-593, S1DGrid.cs:78-147 (host FP64)
//   generator + first step  CLVModel.cs:86-167 / CModel Device.cs:186-196,
//                           here fused and evaluated in FP64 on the GPU
//                           (ophip_sek_rs)
//   repeated squaring       device_mc_sfex, CModel Device.cs:216-228
//   propagation chain       device_mc_smk, CModel Device.cs:269-279
//   pricing                 GEMV against payoff vectors (OPPricer.cs:143-146
//                           discounting), or backward induction with the
//                           skinny GEMM (the role of sgemv3/sgemv4)
#include "ophip.h"
#include "opmodel.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Buf {
    unsigned h = 0;
    size_t bytes = 0;
};

bool alloc(Buf &b, size_t bytes) {
    if (bytes >= (size_t)4 << 30) return false;
    b.h = opcuda_mem_alloc((unsigned)bytes);
    b.bytes = bytes;
    return b.h != 0;
}
void release(Buf &b) {
    if (b.h) opcuda_mem_free_device(b.h);
    b.h = 0;
}
int upload(Buf &b, const void *p, size_t bytes) {
    return opcuda_memcpy_h2d(b.h, (unsigned long)(uintptr_t)p, (unsigned)bytes);
}
// a null result whose cause stays in the library's status: err (the status
// a failing call left, read before any cleanup), or BAD_ARG
template <class T> T *failed(unsigned err) {
    ophip_set_status(err ? err : OPHIP_ERR_BAD_ARG);
    return nullptr;
}

void make_grid(const opm_config *c, std::vector<double> &x, int &x0) {
    const int nx = c->nx;
    x.resize(nx);
    const double h = (c->xhi - c->xlo) / (double)(nx - 1);
    for (int i = 0; i < nx; i++) x[i] = c->xlo + h * (double)i;
    // node nearest to spot, first minimum wins (SGrid.cs:541-551)
    int best = 0;
    double minerr = std::fabs(x[0] - c->spot);
    for (int i = 1; i < nx; i++) {
        double e = std::fabs(x[i] - c->spot);
        if (e < minerr) { best = i; minerr = e; }
    }
    const double scale = c->spot / x[best];  // SGrid.cs:553-557
    for (int i = 0; i < nx; i++) x[i] *= scale;
    x0 = best;
}

double local_vol(const opm_config *c, double x, int r) {
    switch (c->vol_kind) {
    case OPM_VOL_C1: {
        double s = (x > 0.0) ? 0.25 * std::pow(x / 100.0, -0.5) : 1.0;
        return s < 0.05 ? 0.05 : (s > 1.0 ? 1.0 : s);
    }
    case OPM_VOL_CEV: {
        if (x <= 0.0) return 0.0;
        double l = std::log(x / 100.0);
        double s = c->sigma0 * std::exp((c->beta - 1.0) * l) * (1.0 + c->skew * l + c->curv * l * l);
        return s < 1e-4 ? 1e-4 : s;
    }
    default:
        return c->sigma_r[r];
    }
}

void make_params(const opm_config *c, const std::vector<double> &x, std::vector<double> &drift,
                 std::vector<double> &vol) {
    const int nx = c->nx, nr = c->nr;
    drift.assign((size_t)nx * nr, 0.0);
    vol.assign((size_t)nx * nr, 0.0);
    for (int r = 0; r < nr; r++)
        for (int i = 0; i < nx; i++) {
            drift[i + (size_t)nx * r] = c->rate * x[i];          // mu(x) = r x
            vol[i + (size_t)nx * r] = local_vol(c, x[i], r) * x[i];  // SVol in price units
        }
}

bool valid(const opm_config *c) {
    return c && c->nx >= 3 && c->nr >= 1 && c->nr <= 16 && c->xhi > c->xlo && c->DeltaT > 0 &&
           c->nsq >= 0 && c->nsq <= 40 && c->nsteps >= 1 && c->nstrikes >= 1 &&
           c->nstrikes <= 64 && ((c->nx * c->nr) % 64) == 0 && c->precision >= OPM_PREC_FP32 &&
           c->precision <= OPM_PREC_F8X3;
}

}  // namespace

struct opm_solver {
    opm_config cfg;
    int d = 0, x0 = 0, n_mat = 1, ncp = 16;
    double dt0 = 0.0;
    Buf x, drift, vol;   // FP64 lattice inputs
    Buf U[2];            // FP32 d x d ping-pong for the squarings
    Buf V[2], pay0;      // backward induction: d x ncp row-major
    Buf VT[2];           // W form (x3 modes): hi/lo copies of V for the matrix-core step
    Buf payXT;           // the same copies of pay0 (made once: a group's first step reads them)
    // W form, propagation: the spot rows' forward chain p_k = p_{k-1} (I + W)
    Buf fP, fx, fout;
    unsigned long fplan = 0;
    Buf Uk[2], payT, out;  // propagation: U_k, payoff (nstrikes x d), prices
    int u_final = 0, v_final = 0;
    unsigned long graph = 0;
    // U = I + W form (precision > 0): squaring state and plan
    int wmode = -1;
    Buf wstate;
    unsigned long wplan = 0;
    unsigned kernel_h = 0;  // handle of U (plain) or W (W form) after the squarings
    // state order on the device: the W form keeps node-major states y =
    // nr*x + r (one band: its zero-cell skipping applies), the plain method
    // the reference's x + nx*r (S2DGrid.cs:89); prices and opm_kernel are
    // in the reference's order either way
    bool nm = false;
    int state(int x, int r) const { return nm ? x * cfg.nr + r : x + cfg.nx * r; }
    int node(int y) const { return nm ? y / cfg.nr : y % cfg.nx; }
    // the FP16 x3 matrix-core backward steps in the reference's order (its
    // FP32 accumulation is more accurate there): W's fragment copies are
    // gathered through perm (reference state -> device state), and the
    // value vectors V are in the reference's order
    bool v_ref = false;
    Buf perm;
    int vstate(int x, int r) const { return v_ref ? x + cfg.nx * r : state(x, r); }
    int vnode(int y) const { return v_ref ? y % cfg.nx : node(y); }
};

extern "C" {

int opm_grid(const opm_config *cfg, double *xval, int *x0) {
    if (!cfg || cfg->nx < 3) return 1;
    std::vector<double> x;
    int b;
    make_grid(cfg, x, b);
    std::memcpy(xval, x.data(), sizeof(double) * x.size());
    *x0 = b;
    return 0;
}

int opm_params(const opm_config *cfg, double *drift, double *vol) {
    if (!cfg || cfg->nx < 3) return 1;
    std::vector<double> x, dr, vo;
    int b;
    make_grid(cfg, x, b);
    make_params(cfg, x, dr, vo);
    std::memcpy(drift, dr.data(), sizeof(double) * dr.size());
    std::memcpy(vol, vo.data(), sizeof(double) * vo.size());
    return 0;
}

double opm_max_rate(const opm_config *c) {
    if (!c || c->nx < 3) return 0.0;
    std::vector<double> x, dr, vo;
    int b;
    make_grid(c, x, b);
    make_params(c, x, dr, vo);
    const int nx = c->nx;
    double m = 0.0;
    for (int r = 0; r < c->nr; r++)
        for (int i = 1; i < nx - 1; i++) {
            // the two stencil rates of S1DGrid.cs:78-147 / CLVModel.cs:99-114
            const double m00 = x[i - 1] - x[i], m01 = x[i + 1] - x[i];
            const double m10 = m00 * m00, m11 = m01 * m01, det = m00 * m11 - m01 * m10;
            const double mu = dr[i + (size_t)nx * r], v = vo[i + (size_t)nx * r];
            const double am = (m11 * mu - m01 * v * v) / det;
            const double ap = (-m10 * mu + m00 * v * v) / det;
            const double out = am + ap + (c->nr - 1) * c->q_off;
            m = out > m ? out : m;
        }
    return m;
}

int opm_plan_nsq(const opm_config *c, double dt0_cap) {
    const double md = opm_max_rate(c);
    double dt = md > 0.0 ? 0.5 / md : dt0_cap;
    if (dt > dt0_cap) dt = dt0_cap;
    return (int)std::ceil(std::log(c->DeltaT / dt) / std::log(2.0));
}

int opm_sv_params(int nx, int nr, const double *xval, const double *rval, int y0, int ni, const double *t_years,
                  const double *tau, const double *rho, double *svol, double *volvol, double *voldrift,
                  double *jm, double *jp) {
    if (nx < 2 || nr < 1 || ni < 1 || !xval || !rval || !t_years || !tau || !svol || !volvol || !voldrift ||
        !jm || !jp || y0 < 0 || y0 >= nx * nr)
        return 1;
    const int d = nx * nr;
    // per-state price and regime value, state y = x + nx r
    auto xs = [&](int y) { return xval[y % nx]; };
    auto rs = [&](int y) { return rval[y / nx]; };
    const double sx_max = std::sqrt(xs(nx - 1)), sx_min = std::sqrt(xs(0)), x_spot = xs(y0);
    if (!(sx_max > sx_min) || !(x_spot > 0.0)) return 1;
    for (int i = 0; i < ni; i++) {
        // the eight curves at t_i: A + (B - A)(1 + c3 t / c2) exp(-t / c2), or A when c2 = 0,
        // plus the date's offset
        double v[8];
        const double t = t_years[i];
        for (int c = 0; c < 8; c++) {
            const double *q = tau + 4 * c;
            double xi = q[0];
            if (q[2] != 0.0) xi = q[0] + (q[1] - q[0]) * (1.0 + q[3] * t / q[2]) * std::exp(-t / q[2]);
            if (rho) xi += rho[8 * i + c];
            if (!std::isfinite(xi)) return 1;
            v[c] = xi;
        }
        const double vol = v[0], lowbeta = v[1], highbeta = v[2], vv = v[3], volmrr = v[4], volmrl = v[5];
        for (int y = 0; y < d; y++) {
            const double sx = std::sqrt(xs(y));
            const double beta = (highbeta * (sx - sx_min) + lowbeta * (sx_max - sx)) / (sx_max - sx_min);
            const size_t o = (size_t)i * d + y;
            svol[o] = rs(y) * vol * std::pow(xs(y) / x_spot, beta);
            volvol[o] = vv;
            voldrift[o] = volmrr * (volmrl - rs(y));
            jm[o] = v[6];
            jp[o] = v[7];
        }
    }
    return 0;
}

opm_solver *opm_create(const opm_config *cfg) {
    if (!valid(cfg)) return failed<opm_solver>(OPHIP_ERR_BAD_ARG);
    if (opcuda_cublas_init() != 0) return failed<opm_solver>(opcuda_get_status());
    opm_solver *s = new opm_solver();
    s->cfg = *cfg;
    const opm_config *c = &s->cfg;
    const int nx = c->nx, nr = c->nr, d = nx * nr;
    s->d = d;
    s->dt0 = std::ldexp(c->DeltaT, -c->nsq);
    s->ncp = ((c->nstrikes + 15) / 16) * 16;
    std::vector<double> x, dr, vo;
    make_grid(c, x, s->x0);
    make_params(c, x, dr, vo);
    const size_t dd = (size_t)d * d;
    bool ok = alloc(s->x, sizeof(double) * nx) && alloc(s->drift, sizeof(double) * d) &&
              alloc(s->vol, sizeof(double) * d);
    if (c->precision == OPM_PREC_FP32) {
        ok = ok && alloc(s->U[0], 4 * dd) && alloc(s->U[1], 4 * dd);
    } else {
        s->wmode = c->precision - 1;  // OPM_PREC_F16.. -> wsq modes 0..2, 3 = FP8
        s->nm = nr > 1;
        unsigned long wb = s->wmode == 3 ? ophip_w8_bytes(d) : ophip_wsq_bytes(d, s->wmode);
        ok = ok && wb && (d % (s->wmode == 3 ? 64 : 128)) == 0 && alloc(s->wstate, wb);
        if (ok && s->wmode != 3) {
            unsigned st = s->wstate.h;
            s->wplan = ophip_wsq_plan(1, &st, &d, nullptr, s->wmode, 0, 0);
            ok = s->wplan != 0;
        }
    }
    ok = ok && !upload(s->x, x.data(), s->x.bytes) && !upload(s->drift, dr.data(), s->drift.bytes) &&
         !upload(s->vol, vo.data(), s->vol.bytes);
    if (ok && c->price_kind == OPM_PRICE_BACKWARD) {
        s->n_mat = 1;
        size_t vb = 4 * (size_t)d * s->ncp;
        ok = alloc(s->V[0], vb) && alloc(s->V[1], vb) && alloc(s->pay0, vb);
        if (ok && s->wmode == 1) ok = alloc(s->VT[0], vb) && alloc(s->VT[1], vb) && alloc(s->payXT, vb);
        if (ok && s->wmode == 1 && s->nm) {
            s->v_ref = true;
            std::vector<int> pm((size_t)d);
            for (int r = 0; r < nr; r++)
                for (int i = 0; i < nx; i++) pm[(size_t)i + (size_t)nx * r] = s->state(i, r);
            ok = alloc(s->perm, 4 * (size_t)d) && !upload(s->perm, pm.data(), 4 * (size_t)d);
        }
        if (ok) {
            std::vector<float> p((size_t)d * s->ncp, 0.0f);
            for (int y = 0; y < d; y++)
                for (int k = 0; k < c->nstrikes; k++) {
                    double v = x[s->vnode(y)] - c->strikes[k];
                    p[(size_t)y * s->ncp + k] = (float)(v > 0.0 ? v : 0.0);
                }
            ok = !upload(s->pay0, p.data(), vb);
            if (ok && s->wmode == 1 && d % 128 == 0 && d <= 4096)  // sizes the matrix-core step takes
                ok = ophip_wsq_bwd_prep(d, s->ncp, s->wmode, s->pay0.h, s->payXT.h) == 0;
        }
    } else if (ok && s->wmode >= 0) {
        // W form: the forward chain of each regime's spot row, p_k =
        // p_{k-1} (I + W), and its prices (batch.hip), instead of forming
        // every W_k (only the spot rows are priced)
        s->n_mat = c->nsteps;
        const size_t prow = (size_t)(c->nsteps + 1) * d;
        ok = alloc(s->fP, 4 * prow * nr) && alloc(s->fx, 4 * (size_t)d) &&
             alloc(s->fout, 4 * (size_t)nr * c->nsteps * c->nstrikes);
        if (ok) {
            std::vector<float> P0(prow * nr, 0.0f), xs((size_t)d);
            for (int r = 0; r < nr; r++) P0[prow * r + s->state(s->x0, r)] = 1.0f;
            for (int y = 0; y < d; y++) xs[y] = (float)x[s->node(y)];
            ok = !upload(s->fP, P0.data(), 4 * prow * nr) && !upload(s->fx, xs.data(), 4 * (size_t)d);
        }
        if (ok) {
            const unsigned W = s->wmode == 3 ? ophip_w8_w(s->wstate.h, d, c->nsq & 1)
                                             : ophip_wsq_w(s->wstate.h, d, s->wmode, c->nsq & 1);
            std::vector<int> ds(nr, d), plus(nr, 1);
            std::vector<unsigned> K(nr, W), Ph(nr), xh(nr, s->fx.h), oh(nr);
            std::vector<float> rate(nr, (float)c->rate), ks(c->strikes, c->strikes + c->nstrikes);
            for (int r = 0; r < nr; r++) {
                Ph[r] = s->fP.h + (unsigned)(4 * prow * r);
                oh[r] = s->fout.h + (unsigned)(4 * (size_t)r * c->nsteps * c->nstrikes);
            }
            s->fplan = ophip_fwd_plan(nr, ds.data(), K.data(), plus.data(), Ph.data(), xh.data(),
                                      oh.data(), rate.data(), (float)c->DeltaT, c->nsteps, ks.data(),
                                      c->nstrikes);
            ok = s->fplan != 0;
            if (ok && s->wmode != 3) {  // the forward steps skip W's zero cells
                std::vector<unsigned> fl(nr, ophip_wsq_array(s->wstate.h, d, s->wmode, c->nsq & 1, 6));
                ok = ophip_fwd_set_cells(s->fplan, fl.data()) == 0;
            }
        }
    } else if (ok) {
        s->n_mat = c->nsteps;
        ok = alloc(s->Uk[0], 4 * dd) && alloc(s->Uk[1], 4 * dd) &&
             alloc(s->payT, 4 * (size_t)c->nstrikes * d) &&
             alloc(s->out, 4 * (size_t)c->nsteps * nr * c->nstrikes);
        if (ok) {
            // payoff matrix P (nstrikes x d, column-major, ld = nstrikes)
            std::vector<float> p((size_t)c->nstrikes * d);
            for (int y = 0; y < d; y++)
                for (int k = 0; k < c->nstrikes; k++) {
                    double v = x[y % nx] - c->strikes[k];
                    p[k + (size_t)c->nstrikes * y] = (float)(v > 0.0 ? v : 0.0);
                }
            ok = !upload(s->payT, p.data(), s->payT.bytes);
        }
    }
    if (!ok) {
        const unsigned err = opcuda_get_status();
        opm_destroy(s);
        return failed<opm_solver>(err);
    }
    return s;
}

void opm_destroy(opm_solver *s) {
    if (!s) return;
    if (s->graph) ophip_graph_destroy(s->graph);
    if (s->wplan) ophip_wsq_plan_free(s->wplan);
    if (s->fplan) ophip_fwd_plan_free(s->fplan);
    Buf *all[] = {&s->x, &s->drift, &s->vol, &s->U[0], &s->U[1], &s->V[0], &s->V[1],
                  &s->VT[0], &s->VT[1], &s->payXT, &s->pay0, &s->Uk[0], &s->Uk[1], &s->payT, &s->out,
                  &s->wstate, &s->fP, &s->fx, &s->fout, &s->perm};
    for (Buf *b : all) release(*b);
    delete s;
}

// W's fragment layout for the FP16 x3 backward steps: in the reference's
// state order when the squarings ran node-major (the regime permutation,
// through LDS tiles when its shape allows, else gathered through perm)
static int bwd_layout(opm_solver *s, int d, int buf) {
    if (!s->v_ref) return ophip_wsq_bwd_layout(s->wstate.h, d, s->wmode, buf);
    const int nx = s->cfg.nx, nr = s->cfg.nr;
    if (nx % 8 == 0 && (nr == 1 || nr == 2 || nr == 4 || nr == 8))
        return ophip_wsq_bwd_layout_regime(s->wstate.h, d, s->wmode, buf, nx, nr);
    return ophip_wsq_bwd_layout_perm(s->wstate.h, d, s->wmode, buf, s->perm.h);
}

// the solve body: only kernel launches on the library stream (capturable)
static int solve_body(opm_solver *s) {
    const opm_config *c = &s->cfg;
    const int d = s->d, nx = c->nx, nr = c->nr;
    const bool wform = s->wmode >= 0;
    int st;
    if (!wform) {
        st = ophip_sek_rs(nx, nr, s->x.h, s->drift.h, s->vol.h, c->q_off, s->dt0, 0, s->U[0].h);
        if (st) return st;
        // repeated squaring U(2dt) = U(dt)^2, ping-pong (CModel Device.cs:216-228)
        for (int r = 0; r < c->nsq; r++) {
            const Buf &a = s->U[r & 1];
            const Buf &o = s->U[(r + 1) & 1];
            st = ophip_sgemm_strided_batched(d, d, d, 1.0f, a.h, d, 0, a.h, d, 0, 0.0f, o.h, d, 0, 1);
            if (st) return st;
        }
        s->u_final = c->nsq & 1;
        s->kernel_h = s->U[s->u_final].h;
    } else {
        // W0 = dt*L in FP32 straight from the FP64 assembly (fused with
        // the first split in the 16-bit modes), then the W-form squarings
        // with renormalisation
        const bool f8 = s->wmode == 3;
        if (f8) {
            st = ophip_sek_rs(nx, nr, s->x.h, s->drift.h, s->vol.h, c->q_off, s->dt0, s->nm ? 6 : 2,
                              ophip_w8_w(s->wstate.h, d, 0));
            if (st) return st;
            st = ophip_w8_exec(s->wstate.h, d, c->nsq);
        } else {
            const unsigned h = s->wstate.h;
            const double q = c->q_off;
            st = ophip_wsq_assemble_rs(1, &nx, &nr, &s->x.h, &s->drift.h, &s->vol.h, &q, &s->dt0,
                                       s->nm ? 1 : 0, s->wmode, &h);
            if (st) return st;
            st = ophip_wsq_exec_ex(s->wplan, c->nsq, 1);
        }
        if (st) return st;
        s->u_final = c->nsq & 1;
        s->kernel_h = f8 ? ophip_w8_w(s->wstate.h, d, s->u_final)
                         : ophip_wsq_w(s->wstate.h, d, s->wmode, s->u_final);
    }
    const unsigned U = s->kernel_h;
    if (c->price_kind == OPM_PRICE_BACKWARD) {
        const float df = (float)std::exp(-c->rate * c->DeltaT);
        st = opcuda_scopy1(s->V[0].h, s->pay0.h, (unsigned)((size_t)d * s->ncp));
        if (st) return st;
        if (wform && s->wmode == 1) {
            // V <- df*(V + W*V) on the FP16 matrix cores from W's hi/lo
            // copies (a BF16 hi/lo pair of V keeps 16 bits, too few over 252
            // steps: the BF16 method steps in FP32)
            st = bwd_layout(s, d, s->u_final);
            if (st) return st;
            st = ophip_wsq_bwd_prep(d, s->ncp, s->wmode, s->V[0].h, s->VT[0].h);
            if (st) return st;
            // the steps carry the values as their FP16 hi/lo copies; the
            // last step also writes the FP32 values
            for (int t = 0; t < c->nsteps; t++) {
                const int a = t & 1, b = (t + 1) & 1;
                st = ophip_wsq_bwd_step(s->wstate.h, d, s->wmode, s->u_final, s->ncp, df, 0u, s->VT[a].h,
                                        t == c->nsteps - 1 ? s->V[b].h : 0u, s->VT[b].h);
                if (st) return st;
            }
        } else {
            for (int t = 0; t < c->nsteps; t++) {
                // V <- df*U*V, or df*(V + W*V) in the W form
                st = ophip_sgemm_skinny_ex(d, d, s->ncp, df, U, d, s->V[t & 1].h, s->ncp,
                                           s->V[(t + 1) & 1].h, s->ncp, wform ? 1 : 0);
                if (st) return st;
            }
        }
        s->v_final = c->nsteps & 1;
    } else if (s->fplan) {
        st = ophip_fwd_exec(s->fplan);
        if (st) return st;
    } else {
        // U_1 = U; U_k = U_{k-1} * U (CModel Device.cs:269-279); in the W
        // form W_k = W_{k-1} + W + W_{k-1} * W
        unsigned prev = U;
        for (int k = 1; k <= c->nsteps; k++) {
            unsigned cur = U;
            if (k > 1) {
                cur = s->Uk[k & 1].h;
                float beta = 0.0f;
                if (wform) {
                    st = ophip_saxpby(d * d, 1.0f, prev, 1.0f, U, cur);
                    if (st) return st;
                    beta = 1.0f;
                }
                opcuda_sgemm(d, d, d, 1.0f, prev, d, U, d, beta, cur, d);
                st = (int)opcuda_get_status();
                if (st) return st;
            }
            const float df = (float)std::exp(-c->rate * c->DeltaT * k);
            // prices at the spot row of every regime (the payoff GEMV of
            // each row, OPPricer.cs:143-146), one launch per maturity
            std::vector<int> rows(nr);
            for (int r = 0; r < nr; r++) rows[r] = s->x0 + nx * r;
            st = ophip_spot_prices(d, cur, nr, rows.data(), wform ? 1 : 0, s->payT.h, c->nstrikes, df,
                                   s->out.h + 4u * (unsigned)((k - 1) * nr * c->nstrikes));
            if (st) return st;
            prev = cur;
        }
    }
    return 0;
}

int opm_solve(opm_solver *s) {
    if (!s) return 1;
    return solve_body(s);
}

int opm_capture(opm_solver *s) {
    if (!s) return 1;
    if (s->graph) {
        ophip_graph_destroy(s->graph);
        s->graph = 0;
    }
    // size the workspace before capture so that nothing allocates inside
    ophip_reserve_workspace((unsigned long)64 << 20);
    int st = ophip_capture_begin();
    if (st) return st;
    int body = solve_body(s);
    unsigned long g = ophip_capture_end();
    if (body) {
        if (g) ophip_graph_destroy(g);
        return body;
    }
    if (!g) return (int)opcuda_get_status();
    s->graph = g;
    return 0;
}

int opm_launch(opm_solver *s) {
    if (!s || !s->graph) return 1;
    return ophip_graph_launch(s->graph);
}

int opm_prices(opm_solver *s, double *prices) {
    if (!s) return 1;
    opcuda_thread_synchronize();
    const opm_config *c = &s->cfg;
    const int d = s->d, nr = c->nr;
    if (c->price_kind == OPM_PRICE_BACKWARD) {
        std::vector<float> v((size_t)d * s->ncp);
        int st = opcuda_memcpy_d2h(s->V[s->v_final].h, (unsigned long)(uintptr_t)v.data(),
                                   (unsigned)(4 * v.size()));
        if (st) return st;
        for (int r = 0; r < nr; r++)
            for (int k = 0; k < c->nstrikes; k++)
                prices[(size_t)r * c->nstrikes + k] = v[(size_t)s->vstate(s->x0, r) * s->ncp + k];
    } else if (s->fplan) {
        // forward chain: device order [r][k][strike], the output's
        std::vector<float> o((size_t)c->nsteps * nr * c->nstrikes);
        int st = opcuda_memcpy_d2h(s->fout.h, (unsigned long)(uintptr_t)o.data(),
                                   (unsigned)(4 * o.size()));
        if (st) return st;
        for (size_t e = 0; e < o.size(); e++) prices[e] = o[e];
    } else {
        std::vector<float> o((size_t)c->nsteps * nr * c->nstrikes);
        int st = opcuda_memcpy_d2h(s->out.h, (unsigned long)(uintptr_t)o.data(),
                                   (unsigned)(4 * o.size()));
        if (st) return st;
        // device order [k][r][strike] -> output [r][k][strike]
        for (int k = 0; k < c->nsteps; k++)
            for (int r = 0; r < nr; r++)
                for (int j = 0; j < c->nstrikes; j++)
                    prices[((size_t)r * c->nsteps + k) * c->nstrikes + j] =
                        o[((size_t)k * nr + r) * c->nstrikes + j];
    }
    return 0;
}

// the squared kernel (column-major d x d) in the reference's state order
static int kernel_canonical(opm_solver *s, float *K) {
    const int d = s->d;
    const size_t dd = (size_t)d * d;
    if (!s->nm)
        return opcuda_memcpy_d2h(s->kernel_h, (unsigned long)(uintptr_t)K, (unsigned)(4 * dd));
    std::vector<float> t(dd);
    int st = opcuda_memcpy_d2h(s->kernel_h, (unsigned long)(uintptr_t)t.data(), (unsigned)(4 * dd));
    if (st) return st;
    const int nx = s->cfg.nx;
    std::vector<int> canon((size_t)d);  // device state -> reference state
    for (int y = 0; y < d; y++) canon[y] = s->node(y) + nx * (y % s->cfg.nr);
    for (int j = 0; j < d; j++)
        for (int i = 0; i < d; i++) K[(size_t)canon[i] + (size_t)d * canon[j]] = t[(size_t)i + (size_t)d * j];
    return 0;
}

// U(DeltaT) in FP32; in the W form U = I + W is formed on the host
int opm_kernel(opm_solver *s, float *U) {
    if (!s || !s->kernel_h) return 1;
    opcuda_thread_synchronize();
    const int d = s->d;
    int st = kernel_canonical(s, U);
    if (st || s->wmode < 0) return st;
    for (int i = 0; i < d; i++) U[(size_t)i * d + i] += 1.0f;
    return 0;
}

// the raw squared kernel (U, or W in the W form), reference state order
int opm_kernel_raw(opm_solver *s, float *K, int *is_w) {
    if (!s || !s->kernel_h) return 1;
    opcuda_thread_synchronize();
    *is_w = s->wmode >= 0;
    return kernel_canonical(s, K);
}

unsigned opm_kernel_handle(opm_solver *s) { return s ? s->kernel_h : 0; }

int opm_dims(opm_solver *s, int *d, int *n_mat) {
    if (!s) return 1;
    *d = s->d;
    *n_mat = s->n_mat;
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------
// A group of solvers priced together: many scenarios of one lattice size,
// each a whole backward-induction solve in the W form.  The squarings of
// all of them run in one persistent launch per round (one plan over every
// state) and, in FP16 x3, each backward step of all of them in one launch
// (ophip_wsq_bwd_step_batch).  Every solver keeps its own buffers and
// prices; the reference solves one model at a time (CModel Device.cs:
// 186-295), which is what opm_solve on each member does.
// ---------------------------------------------------------------------
struct opm_group {
    std::vector<opm_solver *> s;
    unsigned long wplan = 0;
    unsigned long graph = 0;
};

static int group_body(opm_group *g);

extern "C" {

opm_group *opm_group_create(opm_solver *const *solvers, int n) {
    if (!solvers || n <= 0) return failed<opm_group>(OPHIP_ERR_BAD_ARG);
    const opm_solver *a = solvers[0];
    if (!a) return failed<opm_group>(OPHIP_ERR_BAD_ARG);
    for (int i = 0; i < n; i++) {
        const opm_solver *b = solvers[i];
        if (!b || b->wmode < 0 || b->wmode > 2 || b->cfg.price_kind != OPM_PRICE_BACKWARD ||
            b->d != a->d || b->wmode != a->wmode || b->cfg.nsq != a->cfg.nsq ||
            b->cfg.nsteps != a->cfg.nsteps || b->ncp != a->ncp || b->v_ref != a->v_ref || b->nm != a->nm)
            return failed<opm_group>(OPHIP_ERR_BAD_ARG);
        for (int k = 0; k < i; k++)
            if (solvers[k] == b) return failed<opm_group>(OPHIP_ERR_BAD_ARG);  // each solver once
    }
    opm_group *g = new opm_group();
    g->s.assign(solvers, solvers + n);
    std::vector<unsigned> st(n);
    std::vector<int> ds(n, a->d);
    for (int i = 0; i < n; i++) st[i] = solvers[i]->wstate.h;
    // the squaring plan, picked on the group itself: the tile a single
    // solve of this d uses (ophip_wsq_plan's rule: the same bits) with 1, 2
    // or 4 persistent workgroups per CU, each timed over two whole group
    // solves (zero-cell skipping makes the cost depend on the lattices'
    // bands); the fastest is kept
    const int ncu = opcuda_multi_processor_count(0) > 0 ? opcuda_multi_processor_count(0) : 256;
    const int tile = (long)(a->d / 128) * (a->d / 128) >= ncu ? 128 : 64;
    const int cand[3][2] = {{tile, 2 * ncu}, {tile, ncu}, {tile, 4 * ncu}};
    double best = -1.0;
    unsigned long keep = 0;
    for (const auto &c : cand) {
        g->wplan = ophip_wsq_plan(n, st.data(), ds.data(), nullptr, a->wmode, c[0], c[1]);
        if (!g->wplan) continue;
        double ms = -1.0;
        if (!group_body(g) && !ophip_event_record(14) && !group_body(g) && !group_body(g) &&
            !ophip_event_record(15))
            ms = ophip_event_elapsed_ms(14, 15);
        if (ms >= 0.0 && (best < 0.0 || ms < best)) {
            if (keep) ophip_wsq_plan_free(keep);
            keep = g->wplan;
            best = ms;
        } else {
            ophip_wsq_plan_free(g->wplan);
        }
        g->wplan = 0;
    }
    g->wplan = keep;
    if (!g->wplan) {
        const unsigned err = opcuda_get_status();
        delete g;
        return failed<opm_group>(err);
    }
    return g;
}

void opm_group_destroy(opm_group *g) {
    if (!g) return;
    if (g->graph) ophip_graph_destroy(g->graph);
    if (g->wplan) ophip_wsq_plan_free(g->wplan);
    delete g;
}

}  // extern "C"

// the group's solve body: only kernel launches on the library stream
static int group_body(opm_group *g) {
    const int n = (int)g->s.size();
    opm_solver *a = g->s[0];
    const int d = a->d, wmode = a->wmode, nsq = a->cfg.nsq, nsteps = a->cfg.nsteps;
    int st;
    {  // every scenario's FP64 assembly and first step, fused with the
       // squarings' first split, one launch per 32
        std::vector<int> nx(n), nr(n);
        std::vector<unsigned> xv(n), dr(n), vo(n), states(n);
        std::vector<double> q(n), dt(n);
        for (int i = 0; i < n; i++) {
            const opm_solver *s = g->s[i];
            nx[i] = s->cfg.nx;
            nr[i] = s->cfg.nr;
            xv[i] = s->x.h;
            dr[i] = s->drift.h;
            vo[i] = s->vol.h;
            q[i] = s->cfg.q_off;
            dt[i] = s->dt0;
            states[i] = s->wstate.h;
        }
        st = ophip_wsq_assemble_rs(n, nx.data(), nr.data(), xv.data(), dr.data(), vo.data(), q.data(),
                                   dt.data(), a->nm ? 1 : 0, wmode, states.data());
        if (st) return st;
    }
    st = ophip_wsq_exec_ex(g->wplan, nsq, 1);
    if (st) return st;
    const int buf = nsq & 1;
    std::vector<unsigned> states(n), X[2], XT[2], P0(n), PT0(n);
    std::vector<float> df(n);
    for (int i = 0; i < n; i++) {
        opm_solver *s = g->s[i];
        s->u_final = buf;
        s->kernel_h = ophip_wsq_w(s->wstate.h, d, wmode, buf);
        s->v_final = nsteps & 1;
        if (wmode != 1) {
            st = opcuda_scopy1(s->V[0].h, s->pay0.h, (unsigned)((size_t)d * s->ncp));
            if (st) return st;
        }
        states[i] = s->wstate.h;
        df[i] = (float)std::exp(-s->cfg.rate * s->cfg.DeltaT);
        P0[i] = s->pay0.h;
        PT0[i] = s->payXT.h;
        for (int p = 0; p < 2; p++) {
            X[p].push_back(s->V[p].h);
            XT[p].push_back(s->VT[p].h);
        }
    }
    if (wmode == 1) {
        // FP16 x3 on the matrix cores: every lattice's fragment layout and
        // tile lists in three launches, every lattice's step in one launch;
        // the first step reads the payoff and its copies (made at creation)
        const int nx = a->cfg.nx, nr = a->cfg.nr;
        if (a->v_ref && nx % 8 == 0 && (nr == 1 || nr == 2 || nr == 4 || nr == 8)) {
            st = ophip_wsq_bwd_layout_regime_batch(n, states.data(), d, wmode, buf, nx, nr);
            if (st) return st;
        } else {
            for (opm_solver *s : g->s) {
                st = bwd_layout(s, d, buf);
                if (st) return st;
            }
        }
        // the steps carry the values as their FP16 hi/lo copies (22
        // significant bits); the last step also writes the FP32 values
        const std::vector<unsigned> none(n, 0u);
        for (int t = 0; t < nsteps; t++) {
            const int p = t & 1, q = p ^ 1;
            st = ophip_wsq_bwd_step_batch(n, states.data(), d, wmode, buf, a->ncp, df.data(), none.data(),
                                          t ? XT[p].data() : PT0.data(),
                                          t == nsteps - 1 ? X[q].data() : none.data(), XT[q].data());
            if (st) return st;
        }
        if (nsteps == 0)
            for (opm_solver *s : g->s) {
                st = opcuda_scopy1(s->V[0].h, s->pay0.h, (unsigned)((size_t)d * s->ncp));
                if (st) return st;
            }
    } else {
        for (int t = 0; t < nsteps; t++)
            for (int i = 0; i < n; i++) {
                opm_solver *s = g->s[i];
                st = ophip_sgemm_skinny_ex(d, d, s->ncp, df[i], s->kernel_h, d, s->V[t & 1].h, s->ncp,
                                           s->V[(t + 1) & 1].h, s->ncp, 1);
                if (st) return st;
            }
    }
    return 0;
}

extern "C" {

int opm_group_solve(opm_group *g) { return g ? group_body(g) : 1; }

int opm_group_capture(opm_group *g) {
    if (!g) return 1;
    if (g->graph) {
        ophip_graph_destroy(g->graph);
        g->graph = 0;
    }
    ophip_reserve_workspace((unsigned long)64 << 20);
    int st = ophip_capture_begin();
    if (st) return st;
    int body = group_body(g);
    unsigned long gr = ophip_capture_end();
    if (body) {
        if (gr) ophip_graph_destroy(gr);
        return body;
    }
    if (!gr) return (int)opcuda_get_status();
    g->graph = gr;
    return 0;
}

int opm_group_launch(opm_group *g) { return (g && g->graph) ? ophip_graph_launch(g->graph) : 1; }

}  // extern "C"

// ---------------------------------------------------------------------
// calibration sweep (see opmodel.h)
// ---------------------------------------------------------------------

struct opm_batch {
    int n = 0, nmat = 0, nstrikes = 0, wmode = -1, rounds = 0;
    std::vector<opm_config> cfg;
    std::vector<int> nsq;
    std::vector<int> d, x0;
    std::vector<double> dt0;
    // x, drift, vol, x32, P and out are views into one block `blk`:
    // [drift_0 vol_0 drift_1 vol_1 ...][out_0 out_1 ...][x_i x32_i P_i ...]
    // so that new parameters are one upload and all prices one download
    std::vector<Buf> x, drift, vol, x32, P, out, state, U1;  // U1: plain path ping-pong partner
    Buf blk;
    size_t dv_bytes = 0, out_off = 0;
    unsigned long wplan = 0, fplan = 0, graph = 0;
};

extern "C" {

void opm_batch_destroy(opm_batch *b) {
    if (!b) return;
    if (b->graph) ophip_graph_destroy(b->graph);
    if (b->wplan) ophip_wsq_plan_free(b->wplan);
    if (b->fplan) ophip_fwd_plan_free(b->fplan);
    for (auto *v : {&b->state, &b->U1})
        for (Buf &q : *v) release(q);
    release(b->blk);
    delete b;
}

static int batch_body(opm_batch *b);

// The squaring plan of a batch, picked on the batch itself: the
// autotuner's plan (timed per size bucket with the dense K loop) against
// uniform 128- and 64-tile plans, each timed over two whole solves with
// zero-cell skipping; the fastest is kept.  Runs the solve (outputs are
// overwritten by the caller's next solve).
static bool batch_pick_plan(opm_batch *b) {
    const int n = b->n;
    std::vector<unsigned> st(n);
    for (int i = 0; i < n; i++) st[i] = b->state[i].h;
    std::vector<unsigned long> cand = {b->wplan};
    for (int t : {128, 64}) {
        const unsigned long p = ophip_wsq_plan(n, st.data(), b->d.data(), b->nsq.data(), b->wmode, t, 0);
        if (p) cand.push_back(p);
    }
    double best = -1.0;
    size_t bi = 0;
    for (size_t c = 0; c < cand.size(); c++) {
        b->wplan = cand[c];
        if (batch_body(b)) return false;  // warm
        if (ophip_event_record(14) || batch_body(b) || batch_body(b) || ophip_event_record(15)) return false;
        const double ms = ophip_event_elapsed_ms(14, 15);
        if (ms < 0.0) return false;
        if (best < 0.0 || ms < best) {
            best = ms;
            bi = c;
        }
    }
    for (size_t c = 0; c < cand.size(); c++)
        if (c != bi) ophip_wsq_plan_free(cand[c]);
    b->wplan = cand[bi];
    return true;
}

opm_batch *opm_batch_create(const opm_config *cfgs, int n, int tile) {
    if (!cfgs || n <= 0) return failed<opm_batch>(OPHIP_ERR_BAD_ARG);
    if (opcuda_cublas_init() != 0) return failed<opm_batch>(opcuda_get_status());
    const opm_config &c0 = cfgs[0];
    for (int i = 0; i < n; i++) {
        const opm_config &c = cfgs[i];
        if (!valid(&c) || c.nr != 1 || c.precision != c0.precision ||
            c.nsteps != c0.nsteps || c.DeltaT != c0.DeltaT || c.nstrikes != c0.nstrikes ||
            std::memcmp(c.strikes, c0.strikes, sizeof(double) * c.nstrikes) != 0)
            return failed<opm_batch>(OPHIP_ERR_BAD_ARG);
        if (c.precision != OPM_PREC_FP32 && c.nx % 128) return failed<opm_batch>(OPHIP_ERR_BAD_ARG);
        if (c.precision == OPM_PREC_F8X3) return failed<opm_batch>(OPHIP_ERR_BAD_ARG);  // single lattices only
    }
    opm_batch *b = new opm_batch();
    b->n = n;
    b->nsq.resize(n);
    for (int i = 0; i < n; i++) {
        b->nsq[i] = cfgs[i].nsq;
        b->rounds = std::max(b->rounds, cfgs[i].nsq);
    }
    b->nmat = c0.nsteps;
    b->nstrikes = c0.nstrikes;
    b->wmode = c0.precision == OPM_PREC_FP32 ? -1 : c0.precision - 1;
    b->cfg.assign(cfgs, cfgs + n);
    for (auto *v : {&b->x, &b->drift, &b->vol, &b->x32, &b->P, &b->out, &b->state, &b->U1})
        v->resize(n);
    b->d.resize(n);
    b->x0.resize(n);
    b->dt0.resize(n);
    bool ok = true;
    std::vector<unsigned> Kh(n), Ph(n), xh(n), oh(n);
    std::vector<int> plus(n, b->wmode >= 0 ? 1 : 0);
    std::vector<float> rate(n);
    {  // the block of small arrays
        size_t off = 0;
        for (int i = 0; i < n; i++) off += 16 * (size_t)cfgs[i].nx;
        b->dv_bytes = off;
        off = (off + 255) & ~(size_t)255;
        b->out_off = off;
        off += 4 * (size_t)n * b->nmat * b->nstrikes;
        for (int i = 0; i < n; i++) {
            off = (off + 255) & ~(size_t)255;
            off += 12 * (size_t)cfgs[i].nx + 4 * (size_t)(b->nmat + 1) * cfgs[i].nx + 256;
        }
        ok = alloc(b->blk, off);
        size_t dv = 0, tail = b->out_off + 4 * (size_t)n * b->nmat * b->nstrikes;
        for (int i = 0; i < n && ok; i++) {
            const size_t d = (size_t)cfgs[i].nx;
            auto view = [&](Buf &q, size_t at, size_t bytes) {
                q.h = b->blk.h + (unsigned)at;
                q.bytes = bytes;
            };
            view(b->drift[i], dv, 8 * d);
            view(b->vol[i], dv + 8 * d, 8 * d);
            dv += 16 * d;
            view(b->out[i], b->out_off + 4 * (size_t)i * b->nmat * b->nstrikes,
                 4 * (size_t)b->nmat * b->nstrikes);
            tail = (tail + 255) & ~(size_t)255;
            view(b->x[i], tail, 8 * d);
            view(b->x32[i], tail + 8 * d, 4 * d);
            view(b->P[i], tail + 12 * d + 256 - (12 * d) % 256, 4 * (size_t)(b->nmat + 1) * d);
            tail += 12 * d + 4 * (size_t)(b->nmat + 1) * d + 256;
        }
    }
    for (int i = 0; i < n && ok; i++) {
        const opm_config *c = &b->cfg[i];
        const int d = c->nx;
        b->d[i] = d;
        b->dt0[i] = std::ldexp(c->DeltaT, -c->nsq);
        std::vector<double> x, dr, vo;
        make_grid(c, x, b->x0[i]);
        make_params(c, x, dr, vo);
        const size_t dd = (size_t)d * d;
        if (ok && b->wmode >= 0) ok = alloc(b->state[i], ophip_wsq_bytes(d, b->wmode));
        if (ok && b->wmode < 0) ok = alloc(b->state[i], 4 * dd) && alloc(b->U1[i], 4 * dd);
        if (!ok) break;
        std::vector<float> xf(x.begin(), x.end()), p0((size_t)d, 0.0f);
        p0[b->x0[i]] = 1.0f;  // the chain starts at the spot node
        ok = !upload(b->x[i], x.data(), 8 * (size_t)d) && !upload(b->drift[i], dr.data(), 8 * (size_t)d) &&
             !upload(b->vol[i], vo.data(), 8 * (size_t)d) && !upload(b->x32[i], xf.data(), 4 * (size_t)d) &&
             !upload(b->P[i], p0.data(), 4 * (size_t)d);
        Ph[i] = b->P[i].h;
        xh[i] = b->x32[i].h;
        oh[i] = b->out[i].h;
        rate[i] = (float)c->rate;
        if (b->wmode >= 0) Kh[i] = ophip_wsq_w(b->state[i].h, d, b->wmode, b->nsq[i] & 1);
        else Kh[i] = (b->nsq[i] & 1) ? b->U1[i].h : b->state[i].h;
    }
    if (ok && b->wmode >= 0) {
        std::vector<unsigned> st(n);
        for (int i = 0; i < n; i++) st[i] = b->state[i].h;
        if (tile == 0) {  // the autotuner's tile per size bucket and grid
            std::vector<int> tiles(n, 128);
            int grid = 0;
            ok = ophip_wsq_autotune(n, st.data(), b->d.data(), b->nsq.data(), b->wmode, tiles.data(),
                                    &grid) == 0;
            if (ok)
                b->wplan = ophip_wsq_plan_tiles(n, st.data(), b->d.data(), b->nsq.data(), b->wmode,
                                                tiles.data(), grid);
        } else {
            b->wplan = ophip_wsq_plan(n, st.data(), b->d.data(), b->nsq.data(), b->wmode, tile, 0);
        }
        ok = b->wplan != 0;
    }
    if (ok) {
        std::vector<float> ks(c0.strikes, c0.strikes + c0.nstrikes);
        b->fplan = ophip_fwd_plan(n, b->d.data(), Kh.data(), plus.data(), Ph.data(), xh.data(),
                                  oh.data(), rate.data(), (float)c0.DeltaT, b->nmat, ks.data(),
                                  b->nstrikes);
        ok = b->fplan != 0;
        if (ok && b->wmode >= 0) {  // the forward steps skip W's zero cells
            std::vector<unsigned> fl(n);
            for (int i = 0; i < n; i++) fl[i] = ophip_wsq_array(b->state[i].h, b->d[i], b->wmode, b->nsq[i] & 1, 6);
            ok = ophip_fwd_set_cells(b->fplan, fl.data()) == 0;
        }
    }
    if (ok && b->wmode >= 0 && tile == 0) ok = batch_pick_plan(b);
    if (!ok) {
        const unsigned err = opcuda_get_status();
        opm_batch_destroy(b);
        return failed<opm_batch>(err);
    }
    return b;
}

static int batch_body(opm_batch *b) {
    int st;
    {  // every lattice's FP64 assembly and first step, one launch per 32
       // (W form: fused with the squarings' first split)
        const int n = b->n;
        std::vector<int> nx(n), nr(n, 1);
        std::vector<unsigned> xv(n), dr(n), vo(n), dst(n);
        std::vector<double> q(n, 0.0);
        for (int i = 0; i < n; i++) {
            nx[i] = b->cfg[i].nx;
            xv[i] = b->x[i].h;
            dr[i] = b->drift[i].h;
            vo[i] = b->vol[i].h;
            dst[i] = b->state[i].h;
        }
        if (b->wmode >= 0)
            st = ophip_wsq_assemble_rs(n, nx.data(), nr.data(), xv.data(), dr.data(), vo.data(), q.data(),
                                       b->dt0.data(), 0, b->wmode, dst.data());
        else
            st = ophip_sek_rs_batch(n, nx.data(), nr.data(), xv.data(), dr.data(), vo.data(), q.data(),
                                    b->dt0.data(), 0, dst.data());
        if (st) return st;
    }
    if (b->wmode >= 0) {
        st = ophip_wsq_exec_ex(b->wplan, -1, 1);
        if (st) return st;
    } else {
        for (int r = 0; r < b->rounds; r++)
            for (int i = 0; i < b->n; i++) {
                if (r >= b->nsq[i]) continue;
                const int d = b->d[i];
                const unsigned a = (r & 1) ? b->U1[i].h : b->state[i].h;
                const unsigned o = (r & 1) ? b->state[i].h : b->U1[i].h;
                st = ophip_sgemm_strided_batched(d, d, d, 1.0f, a, d, 0, a, d, 0, 0.0f, o, d, 0, 1);
                if (st) return st;
            }
    }
    return ophip_fwd_exec(b->fplan);
}

int opm_batch_solve(opm_batch *b) { return b ? batch_body(b) : 1; }

int opm_batch_capture(opm_batch *b) {
    if (!b) return 1;
    if (b->graph) {
        ophip_graph_destroy(b->graph);
        b->graph = 0;
    }
    ophip_reserve_workspace((unsigned long)64 << 20);
    int st = ophip_capture_begin();
    if (st) return st;
    int body = batch_body(b);
    unsigned long g = ophip_capture_end();
    if (body) {
        if (g) ophip_graph_destroy(g);
        return body;
    }
    if (!g) return (int)opcuda_get_status();
    b->graph = g;
    return 0;
}

int opm_batch_launch(opm_batch *b) { return (b && b->graph) ? ophip_graph_launch(b->graph) : 1; }

/* new model parameters for every lattice of a batch, keeping its plans
 * and captured graph: each cfgs[i] must match the lattice it replaces in
 * grid (nx, nr, xlo, xhi, spot), squaring count, maturities and strikes;
 * drift, vol and rate may change */
int opm_batch_update(opm_batch *b, const opm_config *cfgs) {
    if (!b || !cfgs) return 1;
    std::vector<float> rate(b->n);
    for (int i = 0; i < b->n; i++) {
        const opm_config &c = cfgs[i], &o = b->cfg[i];
        if (!valid(&c) || c.nx != o.nx || c.nr != o.nr || c.xlo != o.xlo || c.xhi != o.xhi ||
            c.spot != o.spot || c.nsq != o.nsq || c.DeltaT != o.DeltaT || c.nsteps != o.nsteps ||
            c.precision != o.precision || c.nstrikes != o.nstrikes ||
            std::memcmp(c.strikes, o.strikes, sizeof(double) * c.nstrikes) != 0)
            return 0x10003;  /* bad argument (the library's status code) */
    }
    std::vector<double> dv(b->dv_bytes / 8);
    size_t o = 0;
    for (int i = 0; i < b->n; i++) {
        const opm_config *c = &cfgs[i];
        std::vector<double> x, dr, vo;
        int x0 = 0;
        make_grid(c, x, x0);
        make_params(c, x, dr, vo);
        std::copy(dr.begin(), dr.end(), dv.begin() + o);
        std::copy(vo.begin(), vo.end(), dv.begin() + o + dr.size());
        o += dr.size() + vo.size();
        b->cfg[i] = *c;
        rate[i] = (float)c->rate;
    }
    int st = ophip_fwd_set_rates(b->fplan, rate.data());  // waits for the stream first
    if (st) return st;
    return opcuda_memcpy_h2d(b->blk.h, (unsigned long)(uintptr_t)dv.data(), (unsigned)b->dv_bytes);
}

int opm_batch_prices(opm_batch *b, double *prices) {
    if (!b) return 1;
    std::vector<float> o((size_t)b->n * b->nmat * b->nstrikes);
    int st = opcuda_memcpy_d2h(b->blk.h + (unsigned)b->out_off, (unsigned long)(uintptr_t)o.data(),
                               (unsigned)(4 * o.size()));
    if (st) return st;
    for (size_t e = 0; e < o.size(); e++) prices[e] = o[e];
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------
// The step plan of CModel.make_mc_plan (CModel Common.cs:163-644), in days
// as the reference keeps dates (TimeSpan.Days / 365.25), and its device
// execution (device_mc_sfex / sminit / smk / smcopy, CModel Device.cs:
// 186-295) through the drop-in pointer-table ABI (opcuda_sek,
// opcuda_ssqmm, opcuda_scopy1).
//
// Dates: today = day 0; step dates t_k and generator dates t_i strictly
// increasing and > 0; generator i covers (t_{i-1}, t_i].  Merged dates t_q
// (every step and generator date up to the last step) cut time into
// intervals; interval q uses generator i_q.  A kernel j is one (generator,
// interval length) pair, first occurrence first (the reference's todo_q
// dedup, Common.cs:214-264); it starts from I + dt_j L_{i_j} with dt_j =
// DeltaT_j 2^-niter_j, niter_j = ceil(log2(DeltaT_j / dt)), dt = min(0.5 /
// max|L_yy|, dt0) (Common.cs:290-308), and is squared niter_j times.  A step
// kernel m is one step k's product of its intervals' kernels (its chain),
// shared by later steps of the same length inside one generator period
// (todo_k, Common.cs:318-415).
//
// Chain products.  The reference copies the chain's first kernel, then
// multiplies in jfactor_ms[m][s] for s = 0 .. ns-1 (Device.cs:269-279,
// Host Double.cs:227-283): the first factor enters twice and, for the
// longest chain, the last one never does (and shorter chains get extra
// factors).  Here step kernel m is the product of its chain's kernels, each
// once, in date order -- the discretised transition over the step.
// ---------------------------------------------------------------------
struct opm_mcplan {
    int nk = 0, ni = 0, nq = 0, nj = 0, nm = 0, nr = 0, ns = 0;
    std::vector<int> t_q, i_q, j_q, i_j, niter_j, m_k, chain_len;
    std::vector<double> dt_j;
    std::vector<std::vector<int>> chain;  // per m: kernels j in date order
};

extern "C" {

opm_mcplan *opm_mcplan_create(int nk, const int *t_k, int ni, const int *t_i, int dt0_days,
                              const double *maxdiag) {
    if (nk <= 0 || ni <= 0 || !t_k || !t_i || !maxdiag || dt0_days <= 0)
        return failed<opm_mcplan>(OPHIP_ERR_BAD_ARG);
    for (int k = 0; k < nk; k++)
        if (t_k[k] <= 0 || (k > 0 && t_k[k] <= t_k[k - 1])) return failed<opm_mcplan>(OPHIP_ERR_BAD_ARG);
    for (int i = 0; i < ni; i++)
        if (t_i[i] <= 0 || (i > 0 && t_i[i] <= t_i[i - 1]) || !(maxdiag[i] >= 0.0))
            return failed<opm_mcplan>(OPHIP_ERR_BAD_ARG);
    // the last step must lie inside the generators' span (the reference
    // would read a generator past its table)
    if (t_k[nk - 1] > t_i[ni - 1]) return failed<opm_mcplan>(OPHIP_ERR_BAD_ARG);
    opm_mcplan *p = new opm_mcplan();
    p->nk = nk;
    p->ni = ni;
    // merged dates up to the last step
    std::vector<int> all(t_k, t_k + nk);
    all.insert(all.end(), t_i, t_i + ni);
    std::sort(all.begin(), all.end());
    for (int v : all)
        if ((p->t_q.empty() || v > p->t_q.back()) && v <= t_k[nk - 1]) p->t_q.push_back(v);
    p->nq = (int)p->t_q.size();
    // generator of each interval: the first i with t_q <= t_i[i]
    p->i_q.resize(p->nq);
    for (int q = 0, i = 0; q < p->nq; q++) {
        while (p->t_q[q] > t_i[i]) i++;  // t_q[q] <= t_i[ni-1]: stops inside
        p->i_q[q] = i;
    }
    // kernels: distinct (generator, interval length) pairs, first occurrence first
    auto len_q = [&](int q) { return q == 0 ? p->t_q[0] : p->t_q[q] - p->t_q[q - 1]; };
    p->j_q.assign(p->nq, -1);
    for (int q = 0; q < p->nq; q++) {
        for (int q1 = 0; q1 < q; q1++)
            if (p->i_q[q1] == p->i_q[q] && len_q(q1) == len_q(q)) {
                p->j_q[q] = p->j_q[q1];
                break;
            }
        if (p->j_q[q] >= 0) continue;
        p->j_q[q] = p->nj++;
        const int i = p->i_q[q];
        const double DT = len_q(q) / 365.25;
        double dt = maxdiag[i] > 0.0 ? 0.5 / maxdiag[i] : dt0_days / 365.25;
        if (dt > dt0_days / 365.25) dt = dt0_days / 365.25;
        const int niter = (int)std::ceil(std::log(DT / dt) / std::log(2.0));
        p->i_j.push_back(i);
        p->niter_j.push_back(niter);
        p->dt_j.push_back(DT * std::pow(2.0, -niter));
        p->nr = std::max(p->nr, niter);
    }
    // steps: the generator of a step's end and of its start (the first i
    // with t_i[i] >= the date; a step inside one period has both equal)
    auto gen_at = [&](int t) {
        int i = 0;
        while (i < ni && t_i[i] < t) i++;
        return i;
    };
    p->m_k.assign(nk, -1);
    for (int k = 0; k < nk; k++) {
        const int len = k == 0 ? t_k[0] : t_k[k] - t_k[k - 1];
        const int ie = gen_at(t_k[k]), is = k == 0 ? 0 : gen_at(t_k[k - 1]);
        if (ie == is)
            for (int k1 = 0; k1 < k; k1++) {
                const int len1 = k1 == 0 ? t_k[0] : t_k[k1] - t_k[k1 - 1];
                const int ie1 = gen_at(t_k[k1]), is1 = k1 == 0 ? 0 : gen_at(t_k[k1 - 1]);
                if (ie1 == is1 && ie1 == ie && len1 == len) {
                    p->m_k[k] = p->m_k[k1];
                    break;
                }
            }
        if (p->m_k[k] >= 0) continue;
        p->m_k[k] = p->nm++;
        // chain: the intervals inside (t_{k-1}, t_k], in date order
        const int t0 = k == 0 ? 0 : t_k[k - 1], t1 = t_k[k];
        std::vector<int> ch;
        for (int q = 0; q < p->nq; q++) {
            const int a = q == 0 ? 0 : p->t_q[q - 1], b = p->t_q[q];
            if (a >= t0 && b <= t1) ch.push_back(p->j_q[q]);
        }
        p->chain.push_back(ch);
        p->chain_len.push_back((int)ch.size());
        p->ns = std::max(p->ns, (int)ch.size());
    }
    return p;
}

void opm_mcplan_destroy(opm_mcplan *p) { delete p; }

int opm_mcplan_sizes(const opm_mcplan *p, int *nq, int *nj, int *nm, int *nr, int *ns) {
    if (!p) return 1;
    if (nq) *nq = p->nq;
    if (nj) *nj = p->nj;
    if (nm) *nm = p->nm;
    if (nr) *nr = p->nr;
    if (ns) *ns = p->ns;
    return 0;
}

int opm_mcplan_arrays(const opm_mcplan *p, int *t_q, int *i_q, int *j_q, int *i_j, int *niter_j, double *dt_j,
                      int *m_k, int *chain, int *chain_len) {
    if (!p) return 1;
    for (int q = 0; q < p->nq; q++) {
        if (t_q) t_q[q] = p->t_q[q];
        if (i_q) i_q[q] = p->i_q[q];
        if (j_q) j_q[q] = p->j_q[q];
    }
    for (int j = 0; j < p->nj; j++) {
        if (i_j) i_j[j] = p->i_j[j];
        if (niter_j) niter_j[j] = p->niter_j[j];
        if (dt_j) dt_j[j] = p->dt_j[j];
    }
    if (m_k) std::copy(p->m_k.begin(), p->m_k.end(), m_k);
    for (int m = 0; m < p->nm; m++) {
        if (chain_len) chain_len[m] = p->chain_len[m];
        if (chain)
            for (int s = 0; s < p->ns; s++) chain[(size_t)m * p->ns + s] = s < p->chain_len[m] ? p->chain[m][s] : -1;
    }
    return 0;
}

int opm_mcplan_exec(opm_mcplan *p, int d, unsigned gen, unsigned xtpk, unsigned xbuf, unsigned tpk) {
    if (!p || d <= 0 || !gen || !xtpk || !xbuf || !tpk) return 0x10003;
    const unsigned dd = (unsigned)(4 * (size_t)d * d);
    const int nj = p->nj, nm = p->nm;
    // first steps U_j = I + dt_j L_{i_j} (device_mc_sek, Device.cs:186-196)
    for (int j = 0; j < nj; j++) {
        opcuda_sek(d, gen + (unsigned)p->i_j[j] * dd, (float)p->dt_j[j], xtpk + (unsigned)j * dd);
        const int st = (int)opcuda_get_status();
        if (st) return st;
    }
    // device pointer tables: per squaring round the kernels still squaring
    // (ping-pong xtpk_j <-> xbuf_j), per chain position the products
    const int cap = std::max(nj, nm);
    const int rounds = std::max(p->nr, 1), spos = std::max(p->ns - 1, 1);
    const unsigned tab_bytes = (unsigned)(4 * (size_t)cap * (2 * rounds + 3 * spos));
    const unsigned tab = opcuda_mem_alloc(tab_bytes);
    if (!tab) return (int)opcuda_get_status();
    std::vector<unsigned> host(tab_bytes / 4, 0);
    std::vector<int> count_r(rounds, 0), count_s(spos, 0);
    auto A_r = [&](int r) { return (size_t)cap * 2 * r; };
    auto C_r = [&](int r) { return (size_t)cap * (2 * r + 1); };
    auto T_s = [&](int s, int w) { return (size_t)cap * (2 * rounds + 3 * s + w); };  // w: 0 A, 1 B, 2 C
    for (int r = 0; r < p->nr; r++)
        for (int j = 0; j < nj; j++)
            if (r < p->niter_j[j]) {
                const unsigned a = (r & 1) ? xbuf : xtpk, c = (r & 1) ? xtpk : xbuf;
                host[A_r(r) + count_r[r]] = a + (unsigned)j * dd;
                host[C_r(r) + count_r[r]] = c + (unsigned)j * dd;
                count_r[r]++;
            }
    // chain products: position s = 1 .. len-1; odd s: xbuf_m = tpk_m U_j,
    // even s: tpk_m = xbuf_m U_j
    for (int s = 1; s < p->ns; s++)
        for (int m = 0; m < nm; m++)
            if (s < p->chain_len[m]) {
                const unsigned a = (s & 1) ? tpk : xbuf, c = (s & 1) ? xbuf : tpk;
                host[T_s(s - 1, 0) + count_s[s - 1]] = a + (unsigned)m * dd;
                host[T_s(s - 1, 1) + count_s[s - 1]] = xtpk + (unsigned)p->chain[m][s] * dd;
                host[T_s(s - 1, 2) + count_s[s - 1]] = c + (unsigned)m * dd;
                count_s[s - 1]++;
            }
    int st = opcuda_memcpy_h2d(tab, (unsigned long)(uintptr_t)host.data(), tab_bytes);
    // squaring rounds (device_mc_sfex, Device.cs:216-228: A passed twice)
    for (int r = 0; r < p->nr && !st; r++) {
        opcuda_ssqmm(d, count_r[r], tab + 4u * (unsigned)A_r(r), tab + 4u * (unsigned)A_r(r),
                     tab + 4u * (unsigned)C_r(r));
        st = (int)opcuda_get_status();
    }
    // kernels squared an odd number of times end in xbuf (device_mc_scopy)
    for (int j = 0; j < nj && !st; j++)
        if (p->niter_j[j] & 1) st = opcuda_scopy1(xtpk + (unsigned)j * dd, xbuf + (unsigned)j * dd, (unsigned)(d * d));
    // chain heads, products, and the chains of even length end in xbuf
    for (int m = 0; m < nm && !st; m++)
        st = opcuda_scopy1(tpk + (unsigned)m * dd, xtpk + (unsigned)p->chain[m][0] * dd, (unsigned)(d * d));
    for (int s = 0; s + 1 < p->ns && !st; s++) {
        opcuda_ssqmm(d, count_s[s], tab + 4u * (unsigned)T_s(s, 0), tab + 4u * (unsigned)T_s(s, 1),
                     tab + 4u * (unsigned)T_s(s, 2));
        st = (int)opcuda_get_status();
    }
    for (int m = 0; m < nm && !st; m++)
        if (p->chain_len[m] % 2 == 0)
            st = opcuda_scopy1(tpk + (unsigned)m * dd, xbuf + (unsigned)m * dd, (unsigned)(d * d));
    opcuda_thread_synchronize();
    opcuda_mem_free_device(tab);
    return st;
}

}  // extern "C"
