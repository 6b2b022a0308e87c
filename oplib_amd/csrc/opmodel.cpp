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
           c->nstrikes <= 64 && ((c->nx * c->nr) % 16) == 0;
}

}  // namespace

struct opm_solver {
    opm_config cfg;
    int d = 0, x0 = 0, n_mat = 1, ncp = 16;
    double dt0 = 0.0;
    Buf x, drift, vol;   // FP64 lattice inputs
    Buf U[2];            // FP32 d x d ping-pong for the squarings
    Buf V[2], pay0;      // backward induction: d x ncp row-major
    Buf Uk[2], payT, out;  // propagation: U_k, payoff (nstrikes x d), prices
    int u_final = 0, v_final = 0;
    unsigned long graph = 0;
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

opm_solver *opm_create(const opm_config *cfg) {
    if (!valid(cfg)) return nullptr;
    if (opcuda_cublas_init() != 0) return nullptr;
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
              alloc(s->vol, sizeof(double) * d) && alloc(s->U[0], 4 * dd) && alloc(s->U[1], 4 * dd);
    ok = ok && !upload(s->x, x.data(), s->x.bytes) && !upload(s->drift, dr.data(), s->drift.bytes) &&
         !upload(s->vol, vo.data(), s->vol.bytes);
    if (ok && c->price_kind == OPM_PRICE_BACKWARD) {
        s->n_mat = 1;
        size_t vb = 4 * (size_t)d * s->ncp;
        ok = alloc(s->V[0], vb) && alloc(s->V[1], vb) && alloc(s->pay0, vb);
        if (ok) {
            std::vector<float> p((size_t)d * s->ncp, 0.0f);
            for (int y = 0; y < d; y++)
                for (int k = 0; k < c->nstrikes; k++) {
                    double v = x[y % nx] - c->strikes[k];
                    p[(size_t)y * s->ncp + k] = (float)(v > 0.0 ? v : 0.0);
                }
            ok = !upload(s->pay0, p.data(), vb);
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
        opm_destroy(s);
        return nullptr;
    }
    return s;
}

void opm_destroy(opm_solver *s) {
    if (!s) return;
    if (s->graph) ophip_graph_destroy(s->graph);
    Buf *all[] = {&s->x, &s->drift, &s->vol, &s->U[0], &s->U[1], &s->V[0], &s->V[1],
                  &s->pay0, &s->Uk[0], &s->Uk[1], &s->payT, &s->out};
    for (Buf *b : all) release(*b);
    delete s;
}

// the solve body: only kernel launches on the library stream (capturable)
static int solve_body(opm_solver *s) {
    const opm_config *c = &s->cfg;
    const int d = s->d, nx = c->nx, nr = c->nr;
    int st = ophip_sek_rs(nx, nr, s->x.h, s->drift.h, s->vol.h, c->q_off, s->dt0, 0, s->U[0].h);
    if (st) return st;
    // repeated squaring U(2dt) = U(dt)^2, ping-pong (CModel Device.cs:216-228)
    for (int r = 0; r < c->nsq; r++) {
        const Buf &a = s->U[r & 1];
        const Buf &o = s->U[(r + 1) & 1];
        st = ophip_sgemm_strided_batched(d, d, d, 1.0f, a.h, d, 0, a.h, d, 0, 0.0f, o.h, d, 0, 1);
        if (st) return st;
    }
    s->u_final = c->nsq & 1;
    const unsigned U = s->U[s->u_final].h;
    if (c->price_kind == OPM_PRICE_BACKWARD) {
        const float df = (float)std::exp(-c->rate * c->DeltaT);
        st = opcuda_scopy1(s->V[0].h, s->pay0.h, (unsigned)((size_t)d * s->ncp));
        if (st) return st;
        for (int t = 0; t < c->nsteps; t++) {
            st = ophip_sgemm_skinny(d, d, s->ncp, df, U, d, s->V[t & 1].h, s->ncp,
                                    s->V[(t + 1) & 1].h, s->ncp);
            if (st) return st;
        }
        s->v_final = c->nsteps & 1;
    } else {
        // U_1 = U; U_k = U_{k-1} * U (CModel Device.cs:269-279)
        unsigned prev = U;
        for (int k = 1; k <= c->nsteps; k++) {
            unsigned cur = U;
            if (k > 1) {
                cur = s->Uk[k & 1].h;
                st = ophip_sgemm_strided_batched(d, d, d, 1.0f, prev, d, 0, U, d, 0, 0.0f, cur, d,
                                                 0, 1);
                if (st) return st;
            }
            const float df = (float)std::exp(-c->rate * c->DeltaT * k);
            for (int r = 0; r < nr; r++) {
                // row y0 of U_k is the vector at element offset y0, stride d
                const unsigned row = cur + 4u * (unsigned)(s->x0 + nx * r);
                const unsigned dst =
                    s->out.h + 4u * (unsigned)(((k - 1) * nr + r) * c->nstrikes);
                opcuda_sgemv(c->nstrikes, d, df, s->payT.h, c->nstrikes, row, d, 0.0f, dst, 1);
                st = (int)opcuda_get_status();
                if (st) return st;
            }
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
    const int d = s->d, nx = c->nx, nr = c->nr;
    if (c->price_kind == OPM_PRICE_BACKWARD) {
        std::vector<float> v((size_t)d * s->ncp);
        int st = opcuda_memcpy_d2h(s->V[s->v_final].h, (unsigned long)(uintptr_t)v.data(),
                                   (unsigned)(4 * v.size()));
        if (st) return st;
        for (int r = 0; r < nr; r++)
            for (int k = 0; k < c->nstrikes; k++)
                prices[(size_t)r * c->nstrikes + k] = v[(size_t)(s->x0 + nx * r) * s->ncp + k];
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

int opm_kernel(opm_solver *s, float *U) {
    if (!s) return 1;
    opcuda_thread_synchronize();
    return opcuda_memcpy_d2h(s->U[s->u_final].h, (unsigned long)(uintptr_t)U,
                             (unsigned)(4 * (size_t)s->d * s->d));
}

int opm_dims(opm_solver *s, int *d, int *n_mat) {
    if (!s) return 1;
    *d = s->d;
    *n_mat = s->n_mat;
    return 0;
}

}  // extern "C"
