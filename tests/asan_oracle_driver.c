/*
 * Drives every entry point of the CPU oracle (oracle/fastice_oracle.c) on
 * small grids with exactly sized allocations, for a build with
 * -fsanitize=address,undefined (tests/test_oracle_sanitized.py): an access
 * outside a field, a use of freed memory or undefined arithmetic aborts the
 * run.  Test infrastructure only.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fastice_oracle.h"

static unsigned long long rng_state = 0x9e3779b97f4a7c15ull;

static double urand(void)
{
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return (double)(rng_state >> 11) * (1.0 / 9007199254740992.0);
}

static double* field(long n, double lo, double hi)
{
    double* a = (double*)malloc((size_t)n * sizeof(double));
    for (long q = 0; q < n; ++q) a[q] = lo + (hi - lo) * urand();
    return a;
}

static void model_init(or_model* m, int nd, const int n[3], int kind_low, int kind_high)
{
    memset(m, 0, sizeof(*m));
    m->ndims = nd;
    for (int d = 0; d < 3; ++d) {
        m->n[d] = d < nd ? n[d] : 0;
        m->origin[d] = -0.5;
        m->extent[d] = 1.0 + 0.1 * d;
        for (int s = 0; s < 2; ++s) {
            m->bc_kind[d][s] = s == 0 ? kind_low : kind_high;
            for (int c = 0; c < 3; ++c) m->bc_val[d][s][c] = 0.01 * (d + 2 * s + 3 * c);
        }
        m->dV[d] = 1e-3;
        m->grav[d] = d == nd - 1 ? 1.0 : 0.0;
    }
    m->dPr = 0.1;
    for (int q = 0; q < 6; ++q) m->dtau[q] = 0.3;
    m->rh_kind = OR_RHEO_POWERLAW;
    m->rh_A = 1.0;
    m->rh_n = 3.0;
    m->rh_has_reg = 1;
    m->rh_inv_eta_reg = 1.0 / 30.0;
    m->rh_relax = OR_RELAX_LOG;
    m->rh_rate = 0.05;
    or_init_grid(m);
}

static long parent_size(const or_model* m)
{
    long s = 1;
    for (int d = 0; d < m->ndims; ++d) s *= m->n[d] + 3;
    return s;
}

static int run_model(int nd, const int n[3], int ib, int kind_low, int kind_high)
{
    or_model m;
    model_init(&m, nd, n, kind_low, kind_high);
    const long np = parent_size(&m);
    or_fields f;
    memset(&f, 0, sizeof(f));
    double** all[] = {&f.P, &f.txx, &f.tyy, &f.tzz, &f.txy, &f.txz, &f.tyz, &f.Vx, &f.Vy, &f.Vz,
                      &f.rP, &f.rVx, &f.rVy, &f.rVz};
    for (size_t q = 0; q < sizeof(all) / sizeof(all[0]); ++q) *all[q] = field(np, -1e-2, 1e-2);
    f.eta = field(np, 0.5, 2.0);
    f.eta_next = field(np, 0.5, 2.0);
    const int nq = nd == 3 ? 8 : 4;
    if (ib)
        for (int q = 0; q < nq; ++q) {
            f.na[q] = field(np, 0.0, 1.0);
            f.ns[q] = field(np, 0.0, 1.0);
            for (long c = 0; c < np; c += 7) f.na[q][c] = 0.0;
        }
    unsigned char* flags = (unsigned char*)malloc((size_t)np);
    double norms[4];
    or_advance(&m, &f, 7, ib);
    or_compute_residuals(&m, &f, ib);
    or_residual_norms(&m, &f, norms);
    or_compute_tau(&m, &f, ib);
    or_update_sigma(&m, &f, ib);
    or_update_V(&m, &f, ib);
    or_bc_stress(&m, &f);
    or_bc_velocity(&m, &f);
    or_bc_neumann_center(&m, f.eta);
    if (ib) or_ib_flags(&m, &f, flags);
    int it = or_solve(&m, &f, 4, 2, 1e-30, ib, norms);
    double* psi = field(np, -1.0, 1.0);
    double* om[8] = {0};
    for (int q = 0; q < nq; ++q) om[q] = field(np, 0.0, 0.0);
    or_omega_from_psi(&m, psi, om);
    or_invert_levelset(&m, psi);
    or_bc_neumann_vertex(&m, psi);
    if (nd == 3) {
        const int dn[2] = {6, 5};
        double* dem = field((long)(dn[0] + 3) * (dn[1] + 3), 0.1, 0.4);
        const double o[2] = {-0.6, -0.6}, e[2] = {1.2, 1.3};
        const double R[9] = {0.8, -0.6, 0.0, 0.6, 0.8, 0.0, 0.0, 0.0, 1.0};
        or_levelset_from_dem(&m, psi, dem, dn, o, e, NULL);
        or_levelset_from_dem(&m, psi, dem, dn, o, e, R);
        free(dem);
    } else {
        const double poly[8] = {-0.3, -0.2, 0.4, -0.25, 0.35, 0.3, -0.2, 0.25};
        or_levelset_from_polygon(&m, psi, poly, 4);
    }
    for (int q = 0; q < 8; ++q) free(om[q]);
    free(psi);
    free(flags);
    for (size_t q = 0; q < sizeof(all) / sizeof(all[0]); ++q) free(*all[q]);
    free(f.eta);
    free(f.eta_next);
    if (ib)
        for (int q = 0; q < nq; ++q) {
            free(f.na[q]);
            free(f.ns[q]);
        }
    return it;
}

static void run_heat(int nd, const int n[3])
{
    or_heat h;
    memset(&h, 0, sizeof(h));
    h.ndims = nd;
    for (int d = 0; d < 3; ++d) {
        h.n[d] = d < nd ? n[d] : 0;
        h.dx[d] = 0.1;
        h.bc_kind[d][0] = OR_HEAT_VALUE;
        h.bc_kind[d][1] = OR_HEAT_FLUX;
        h.bc_val[d][0] = 1.0;
        h.bc_val[d][1] = 0.5;
    }
    h.sy = n[0] + 3;
    h.sz = nd == 3 ? (long)(n[0] + 3) * (n[1] + 3) : 0;
    h.lam = 1.0;
    h.dtq = 0.01;
    h.dtT = 0.01;
    long np = 1;
    for (int d = 0; d < nd; ++d) np *= n[d] + 3;
    or_heat_fields f;
    memset(&f, 0, sizeof(f));
    f.T = field(np, 0.0, 1.0);
    f.To = field(np, 0.0, 1.0);
    for (int d = 0; d < nd; ++d) f.q[d] = field(np, -0.1, 0.1);
    double out[2];
    or_heat_bc(&h, &f);
    or_heat_advance(&h, &f, 0.1, 5);
    or_heat_residual_norms(&h, &f, 0.1, out);
    free(f.T);
    free(f.To);
    for (int d = 0; d < nd; ++d) free(f.q[d]);
}

int main(void)
{
    const int n3[3] = {6, 5, 4}, n2[3] = {7, 5, 0};
    for (int ib = 0; ib < 2; ++ib) {
        run_model(3, n3, ib, OR_BC_SLIP, OR_BC_TRACTION);
        run_model(3, n3, ib, OR_BC_VELOCITY, OR_BC_SLIP);
        run_model(2, n2, ib, OR_BC_TRACTION, OR_BC_VELOCITY);
    }
    run_heat(3, n3);
    run_heat(2, n2);
    double p8[8], d3[3] = {0.1, 0.2, 0.3}, p4[4], d2[2] = {0.1, 0.2};
    for (int q = 0; q < 8; ++q) p8[q] = urand() - 0.5;
    for (int q = 0; q < 4; ++q) p4[q] = urand() - 0.5;
    printf("omega3 %.6f omega2 %.6f sd %.6f\n", or_psi2omega3(p8, d3), or_psi2omega2(p4, d2),
           or_sd_poly(0.1, 0.1, (const double[]){0, 0, 1, 0, 1, 1, 0, 1}, 4));
    printf("sanitized oracle run complete\n");
    return 0;
}
