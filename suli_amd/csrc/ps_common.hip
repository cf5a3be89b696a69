// Helpers shared by the C-ABI front ends (see ps_common.h).
#include "ps_common.h"

#include <cmath>
#include <cstdio>
#include <cstring>

namespace sps {

namespace {
int bad(std::string &err, const char *fmt, int a = 0, int b = 0, int c = 0)
{
    char buf[256];
    std::snprintf(buf, sizeof buf, fmt, a, b, c);
    err = buf;
    return SULI_PS_EINVAL;
}
}  // namespace

int validate_shape(const suli_ps_grid *g, std::string &err)
{
    if (!g) return bad(err, "grid is NULL");
    if (g->nx < 2 || g->ny < 2 || g->nz < 2)
        return bad(err, "grid %dx%dx%d: every dimension must be >= 2", g->nx, g->ny, g->nz);
    if ((long long)g->nx * g->ny * g->nz > (1LL << 31) - 1)
        return bad(err, "grid %dx%dx%d is too large", g->nx, g->ny, g->nz);
    if (!(g->dx > 0) || !(g->dy > 0) || !(g->dt > 0)) return bad(err, "dx, dy and dt must be positive");
    return 0;
}

int validate_grid(const suli_ps_grid *g, std::string &err)
{
    const int rc = validate_shape(g, err);
    if (rc) return rc;
    if (g->t_press != 1 && g->t_press != 2)
        return bad(err, "t_press = %d: only 1 (projection) and 2 (IPA) solve for pressure", g->t_press);
    if (g->t_press == 2 && !(g->rho_f1 > 0)) return bad(err, "IPA needs rho_f1 > 0");
    if (g->top_bc != 0 && g->top_bc != 1) return bad(err, "top_bc = %d: 0 (zero flux) or 1 (p = 0)", g->top_bc);
    return 0;
}

int validate_problem(const suli_ps_grid *g, const double *u, const double *v, const double *w, const double *rho,
                     const double *ox, const double *oy, const double *oz, const double *dz, const double *dzz,
                     const double *prd1, std::string &err)
{
    int rc = validate_grid(g, err);
    if (rc) return rc;
    if (!u || !v || !w || !prd1 || (g->t_press == 1 && !rho))
        return bad(err, "u, v, w, rho and prd1 must be non-NULL");
    if ((dz == nullptr) != (dzz == nullptr)) return bad(err, "dz and dzz go together");
    const bool mask = g->obst_t != 0 && g->ibm_t == 2;
    if (mask && (!ox || !oy || !oz)) return bad(err, "IBM mask on (obst_t, ibm_t=2) needs obs_ls*");
    if (!dz && !(g->dz > 0)) return bad(err, "dz must be positive");
    return 0;
}

SetupIn make_setup(const suli_ps_grid *g, const double *u, const double *v, const double *w, const double *rho,
                   const double *prd1, const double *ox, const double *oy, const double *oz, const double *dz,
                   const double *dzz, const uint8_t *solid)
{
    SetupIn in{};
    in.solid = solid;
    in.u = u;
    in.v = v;
    in.w = w;
    in.rho = rho;
    in.prd1 = prd1;
    const bool mask = g->obst_t != 0 && g->ibm_t == 2;
    in.obsx = mask ? ox : nullptr;
    in.obsy = mask ? oy : nullptr;
    in.obsz = mask ? oz : nullptr;
    in.dz = dz;
    in.dzz = dzz;
    in.dx = g->dx;
    in.dy = g->dy;
    in.dzu = g->dz;
    in.dt = g->dt;
    // aux = dt / (delta*delta), 9_grad.f90:93-100
    in.aux1 = g->dt / (g->dx * g->dx);
    in.aux2 = g->dt / (g->dy * g->dy);
    in.aux3 = g->dt / (g->dz * g->dz);
    in.rho_f1 = g->rho_f1;
    in.ipa = g->t_press == 2;
    return in;
}

SolveOpts solve_opts(const suli_ps_opts &o)
{
    SolveOpts so;
    so.mixed = o.precision == SULI_PS_MIXED;
    so.stop_rule = o.stop_rule;
    so.tol = o.tol;
    so.max_iter = o.max_iter > 0 ? o.max_iter : 10000;
    so.keep_null = o.keep_null != 0;
    so.store_r = o.stored_residual != 0;
    return so;
}

void fill_stats(const Slot &sl, suli_ps_stats *st)
{
    const CgScalars &f = sl.status;
    std::memset(st, 0, sizeof *st);
    st->status = f.status == 0 ? SULI_PS_OK : SULI_PS_NOT_CONVERGED;
    st->iters = f.iters;
    st->refreshes = f.refreshes;
    st->rr0 = f.rr0;
    st->bb = f.bb;
    // r.r the stop rule saw: FP64, the recurrence's (slot of iteration done_at);
    // mixed, the FP64 true residual of the last replacement
    st->rr = sl.o.mixed ? f.rr_true : f.rr[(f.done_at == kNotDone ? 0 : f.done_at) & 1];
    st->thresh = f.thresh;
    st->rel_true = f.bb > 0 ? std::sqrt(f.rr_true / f.bb) : 0.0;
}

}  // namespace sps
