/*
 * fastice_oracle.c -- CPU FP64 restatement of FastIce.jl's pseudo-transient
 * Stokes hot path, written loop by loop from the reference kernels.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker for libfasticehip.so.  The
 * product path never links, loads or calls it.  Built with
 * -ffp-contract=off so every expression rounds exactly as written; the
 * evaluation order of each expression follows Julia's left-to-right
 * parsing of the reference source line it cites.
 *
 * Semantics the reference delegates to Chmy.jl (unvendored; see the
 * oracle header for the storage convention):
 *   d/dx of a Vertex-in-x field at Center index i: (f[i+1]-f[i]) * (1/dx)
 *   d/dx of a Center-in-x field at Vertex index i: (f[i]-f[i-1]) * (1/dx)
 *   lerp Center->Vertex: 0.5*(f[i-1]+f[i]); Vertex->Center: 0.5*(f[i]+f[i+1]);
 *     multi-dimensional lerps nest x innermost, then y, then z.
 *   masked d/dx(f, w): (w_f[r]*f[r] - w_f[l]*f[l]) * (1/dx), w_f = mask at f's location
 *   kernels run over Chmy indices [1, n+1] in every dimension;
 *   boundary conditions run dimension by dimension (x, y, z), low side then
 *   high side, over the field's full padded extent in the other dimensions:
 *     Dirichlet(v), Vertex in normal dim: boundary node = v
 *     Dirichlet(v), Center in normal dim: ghost = 2v - first interior
 *     Neumann(),   Center: ghost = first interior
 *     Neumann(),   Vertex: ghost = mirror of the node next to the boundary node
 */
#include "fastice_oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#define TOL_EMPTY 1e-6 /* null_spaces.jl:16,41 */

void or_init_grid(or_model* m)
{
    for (int d = 0; d < 3; ++d) {
        if (d < m->ndims) {
            m->dx[d] = m->extent[d] / (double)m->n[d];
            m->idx[d] = 1.0 / m->dx[d];
        } else {
            m->n[d] = 0;
            m->dx[d] = 1.0;
            m->idx[d] = 1.0;
        }
    }
    m->sy = m->n[0] + 3;
    m->sz = (m->ndims == 3) ? (long)(m->n[0] + 3) * (m->n[1] + 3) : 0;
}

/* ------------------------------------------------------------------ */
/* rheology closures: Physics.jl:11,18,26,46-47                        */
/* ------------------------------------------------------------------ */
static double ipow(double x, double p)
{
    /* integer exponents as Julia's x^n (n::Integer): x*x for n=2, x*x*x for n=3 */
    if (p == 0.0) return 1.0;
    if (p == 1.0) return x;
    if (p == 2.0) return x * x;
    if (p == 3.0) return x * x * x;
    return pow(x, p);
}

static double rheo_inner(const or_model* m, const or_fields* f, long c, double tII)
{
    switch (m->rh_kind) {
    case OR_RHEO_LINEAR_CONST:
        return m->rh_eta;                                  /* Physics.jl:11 */
    case OR_RHEO_LINEAR_FIELD:
        return f->rh_field[c];                             /* Physics.jl:11 */
    default: {
        double A = f->rh_field ? f->rh_field[c] : m->rh_A; /* Physics.jl:18 */
        return 0.5 / (A * ipow(tII, m->rh_n - 1.0));
    }
    }
}

double or_rheology(const or_model* m, const or_fields* f, long c, double tII)
{
    double eta = rheo_inner(m, f, c, tII);
    if (m->rh_has_reg) eta = 1.0 / (1.0 / eta + m->rh_inv_eta_reg); /* Physics.jl:26 */
    if (m->rh_relax != OR_RELAX_NONE) {
        /* Physics.jl:46-47; the relaxed field is the lagged viscosity eta */
        double old = f->eta[c];
        double w = 1.0 - m->rh_rate;
        if (m->rh_relax == OR_RELAX_LIN)
            eta = old * w + eta * m->rh_rate;
        else if (m->rh_relax == OR_RELAX_LOG)
            eta = exp(log(old) * w + log(eta) * m->rh_rate);
        else
            eta = pow(10.0, log10(old) * w + log10(eta) * m->rh_rate);
    }
    return eta;
}

/* ------------------------------------------------------------------ */
/* boundary conditions (Chmy bc! semantics as specified above)         */
/* ------------------------------------------------------------------ */
enum { BC_DIRICHLET = 0, BC_NEUMANN = 1 };

/* location bit d set = Vertex in dim d */
enum { L_CCC = 0, L_VCC = 1, L_CVC = 2, L_VVC = 3, L_CCV = 4, L_VCV = 5, L_CVV = 6, L_VVV = 7 };

static void face_apply(const or_model* m, double* a, int L, int d, int s, int kind, double v)
{
    const long st[3] = {1, m->sy, m->sz};
    const int nd = m->ndims;
    int ext[3] = {1, 1, 1};
    for (int e = 0; e < nd; ++e) ext[e] = m->n[e] + ((L >> e) & 1) + 2;
    const int vert = (L >> d) & 1;
    const int nn = m->n[d];
    long tgt, src = -1;
    if (kind == BC_DIRICHLET) {
        if (vert) tgt = (s == 0) ? 1 : nn + 1;
        else { tgt = (s == 0) ? 0 : nn + 1; src = (s == 0) ? 1 : nn; }
    } else {
        if (vert) { tgt = (s == 0) ? 0 : nn + 2; src = (s == 0) ? 2 : nn; }
        else { tgt = (s == 0) ? 0 : nn + 1; src = (s == 0) ? 1 : nn; }
    }
    int e1 = -1, e2 = -1;
    for (int e = 0; e < nd; ++e) {
        if (e == d) continue;
        if (e1 < 0) e1 = e; else e2 = e;
    }
    const int n1 = ext[e1], n2 = (e2 >= 0) ? ext[e2] : 1;
    const long s1 = st[e1], s2 = (e2 >= 0) ? st[e2] : 0;
    for (int q2 = 0; q2 < n2; ++q2)
        for (int q1 = 0; q1 < n1; ++q1) {
            long base = q1 * s1 + q2 * s2;
            long t = base + tgt * st[d];
            if (kind == BC_DIRICHLET) {
                if (vert) a[t] = v;
                else a[t] = 2.0 * v - a[base + src * st[d]];
            } else {
                a[t] = a[base + src * st[d]];
            }
        }
}

void or_bc_neumann_center(const or_model* m, double* a)
{
    for (int d = 0; d < m->ndims; ++d)
        for (int s = 0; s < 2; ++s) face_apply(m, a, L_CCC, d, s, BC_NEUMANN, 0.0);
}

void or_bc_neumann_vertex(const or_model* m, double* a)
{
    int L = (m->ndims == 3) ? L_VVV : L_VVC;
    for (int d = 0; d < m->ndims; ++d)
        for (int s = 0; s < 2; ++s) face_apply(m, a, L, d, s, BC_NEUMANN, 0.0);
}

/* stress batch: boundary_conditions.jl:26-100 */
void or_bc_stress(const or_model* m, or_fields* f)
{
    for (int d = 0; d < m->ndims; ++d)
        for (int s = 0; s < 2; ++s) {
            const int kind = m->bc_kind[d][s];
            const double* c = m->bc_val[d][s];
            if (m->ndims == 3) {
                if (kind == OR_BC_TRACTION) {
                    double* diag = (d == 0) ? f->txx : (d == 1) ? f->tyy : f->tzz;
                    face_apply(m, f->P, L_CCC, d, s, BC_DIRICHLET, c[d]);
                    face_apply(m, diag, L_CCC, d, s, BC_DIRICHLET, c[d]);
                }
                if (kind == OR_BC_TRACTION || kind == OR_BC_SLIP) {
                    if (d == 0) {
                        face_apply(m, f->txy, L_VVC, d, s, BC_DIRICHLET, c[1]);
                        face_apply(m, f->txz, L_VCV, d, s, BC_DIRICHLET, c[2]);
                    } else if (d == 1) {
                        face_apply(m, f->txy, L_VVC, d, s, BC_DIRICHLET, c[0]);
                        face_apply(m, f->tyz, L_CVV, d, s, BC_DIRICHLET, c[2]);
                    } else {
                        face_apply(m, f->txz, L_VCV, d, s, BC_DIRICHLET, c[0]);
                        face_apply(m, f->tyz, L_CVV, d, s, BC_DIRICHLET, c[1]);
                    }
                }
            } else {
                if (kind == OR_BC_TRACTION) {
                    double* diag = (d == 0) ? f->txx : f->tyy;
                    face_apply(m, f->P, L_CCC, d, s, BC_DIRICHLET, c[d]);
                    face_apply(m, diag, L_CCC, d, s, BC_DIRICHLET, c[d]);
                }
                if (kind == OR_BC_TRACTION || kind == OR_BC_SLIP)
                    face_apply(m, f->txy, L_VVC, d, s, BC_DIRICHLET, c[1 - d]);
            }
        }
}

/* velocity batch + eta_next Neumann: boundary_conditions.jl:104-115, Isothermal.jl:103-104 */
void or_bc_velocity(const or_model* m, or_fields* f)
{
    double* V[3] = {f->Vx, f->Vy, f->Vz};
    const int LV[3] = {L_VCC, L_CVC, L_CCV};
    for (int d = 0; d < m->ndims; ++d)
        for (int s = 0; s < 2; ++s) {
            const int kind = m->bc_kind[d][s];
            const double* c = m->bc_val[d][s];
            if (kind == OR_BC_VELOCITY) {
                for (int q = 0; q < m->ndims; ++q) face_apply(m, V[q], LV[q], d, s, BC_DIRICHLET, c[q]);
            } else if (kind == OR_BC_SLIP) {
                face_apply(m, V[d], LV[d], d, s, BC_DIRICHLET, c[d]);
            }
            face_apply(m, f->eta_next, L_CCC, d, s, BC_NEUMANN, 0.0);
        }
}

/* residual batch: boundary_conditions.jl:120-128 */
static void or_bc_residual(const or_model* m, or_fields* f)
{
    double* R[3] = {f->rVx, f->rVy, f->rVz};
    const int LV[3] = {L_VCC, L_CVC, L_CCV};
    for (int d = 0; d < m->ndims; ++d)
        for (int s = 0; s < 2; ++s)
            if (m->bc_kind[d][s] != OR_BC_TRACTION) face_apply(m, R[d], LV[d], d, s, BC_DIRICHLET, 0.0);
}

/* ------------------------------------------------------------------ */
/* null spaces: null_spaces.jl:1-43                                     */
/* ------------------------------------------------------------------ */
#define LT(w, c) ((w)[c] < TOL_EMPTY)

static int null3(double* const* w, int loc, long c, long sy, long sz)
{
    switch (loc) {
    case L_CCC: return LT(w[OM_VCC], c) || LT(w[OM_VCC], c + 1) || LT(w[OM_CVC], c) || LT(w[OM_CVC], c + sy) ||
                       LT(w[OM_CCV], c) || LT(w[OM_CCV], c + sz);
    case L_VCC: return LT(w[OM_CCC], c - 1) || LT(w[OM_CCC], c) || LT(w[OM_VVC], c) || LT(w[OM_VVC], c + sy) ||
                       LT(w[OM_VCV], c) || LT(w[OM_VCV], c + sz);
    case L_CVC: return LT(w[OM_VVC], c) || LT(w[OM_VVC], c + 1) || LT(w[OM_CCC], c - sy) || LT(w[OM_CCC], c) ||
                       LT(w[OM_CVV], c) || LT(w[OM_CVV], c + sz);
    case L_CCV: return LT(w[OM_VCV], c) || LT(w[OM_VCV], c + 1) || LT(w[OM_CVV], c) || LT(w[OM_CVV], c + sy) ||
                       LT(w[OM_CCC], c - sz) || LT(w[OM_CCC], c);
    case L_VVC: return LT(w[OM_CVC], c - 1) || LT(w[OM_CVC], c) || LT(w[OM_VCC], c - sy) || LT(w[OM_VCC], c);
    case L_VCV: return LT(w[OM_CCV], c - 1) || LT(w[OM_CCV], c) || LT(w[OM_VCC], c - sz) || LT(w[OM_VCC], c);
    case L_CVV: return LT(w[OM_CCV], c - sy) || LT(w[OM_CCV], c) || LT(w[OM_CVC], c - sz) || LT(w[OM_CVC], c);
    }
    return 1;
}

static int comp3(int loc)
{
    switch (loc) {
    case L_CCC: return OM_CCC;
    case L_VCC: return OM_VCC;
    case L_CVC: return OM_CVC;
    case L_CCV: return OM_CCV;
    case L_VVC: return OM_VVC;
    case L_VCV: return OM_VCV;
    case L_CVV: return OM_CVV;
    default: return OM_VVV;
    }
}

/* isvalid(w, loc) && !isempty(w2, loc) */
static int valid3(double* const* w, double* const* w2, int loc, long c, long sy, long sz)
{
    int k = comp3(loc);
    return !(null3(w, loc, c, sy, sz) || LT(w[k], c)) && !LT(w2[k], c);
}

/* 2D null spaces (the generic rule of null_spaces.jl:13-22 in two dimensions) */
static int null2(double* const* w, int loc, long c, long sy)
{
    switch (loc) {
    case L_CCC: return LT(w[OM2_VC], c) || LT(w[OM2_VC], c + 1) || LT(w[OM2_CV], c) || LT(w[OM2_CV], c + sy);
    case L_VCC: return LT(w[OM2_CC], c - 1) || LT(w[OM2_CC], c) || LT(w[OM2_VV], c) || LT(w[OM2_VV], c + sy);
    case L_CVC: return LT(w[OM2_VV], c) || LT(w[OM2_VV], c + 1) || LT(w[OM2_CC], c - sy) || LT(w[OM2_CC], c);
    case L_VVC: return LT(w[OM2_CV], c - 1) || LT(w[OM2_CV], c) || LT(w[OM2_VC], c - sy) || LT(w[OM2_VC], c);
    }
    return 1;
}

static int comp2(int loc)
{
    switch (loc) {
    case L_CCC: return OM2_CC;
    case L_VCC: return OM2_VC;
    case L_CVC: return OM2_CV;
    default: return OM2_VV;
    }
}

static int valid2(double* const* w, double* const* w2, int loc, long c, long sy)
{
    int k = comp2(loc);
    return !(null2(w, loc, c, sy) || LT(w[k], c)) && !LT(w2[k], c);
}

/* Validity bits over the work range, the predicates the masked kernels test:
 * stress locations isvalid(na) && !isempty(ns) (IB kernels_3d.jl:6,26,32,38),
 * velocity locations isvalid(ns) && !isempty(na) (:50,56,62).  Bits 0-3:
 * ccc vvc vcv cvv; bits 4-6: vcc cvc ccv (2D: bits 0, 1, 4, 5). */
void or_ib_flags(const or_model* m, const or_fields* f, unsigned char* out)
{
    const int nd = m->ndims;
    const long sy = m->sy, sz = m->sz;
    const long total = (nd == 3) ? sz * (m->n[2] + 3) : sy * (m->n[1] + 3);
    memset(out, 0, (size_t)total);
    const int locs3[7] = {L_CCC, L_VVC, L_VCV, L_CVV, L_VCC, L_CVC, L_CCV};
    const int locs2[4] = {L_CCC, L_VVC, L_VCC, L_CVC};
    const int bit2[4] = {0, 1, 4, 5};
    const int khi = (nd == 3) ? m->n[2] + 1 : 0, klo = (nd == 3) ? 1 : 0;
    for (int k = klo; k <= khi; ++k)
        for (int j = 1; j <= m->n[1] + 1; ++j)
            for (int i = 1; i <= m->n[0] + 1; ++i) {
                const long c = i + sy * j + sz * k;
                unsigned v = 0;
                if (nd == 3) {
                    for (int q = 0; q < 7; ++q) {
                        int ok = (q < 4) ? valid3(f->na, f->ns, locs3[q], c, sy, sz) : valid3(f->ns, f->na, locs3[q], c, sy, sz);
                        if (ok) v |= 1u << q;
                    }
                } else {
                    for (int q = 0; q < 4; ++q) {
                        int ok = (q < 2) ? valid2(f->na, f->ns, locs2[q], c, sy) : valid2(f->ns, f->na, locs2[q], c, sy);
                        if (ok) v |= 1u << bit2[q];
                    }
                }
                out[c] = (unsigned char)v;
            }
}

/* ------------------------------------------------------------------ */
/* 3D kernels: kernels_3d.jl (FullStokes) and IB kernels_3d.jl          */
/* ------------------------------------------------------------------ */
static void sigma3(const or_model* m, or_fields* f, int ib)
{
    const int nx = m->n[0], ny = m->n[1], nz = m->n[2];
    const long sy = m->sy, sz = m->sz;
    const double ix = m->idx[0], iy = m->idx[1], iz = m->idx[2];
    double *P = f->P, *txx = f->txx, *tyy = f->tyy, *tzz = f->tzz, *txy = f->txy, *txz = f->txz, *tyz = f->tyz;
    const double *Vx = f->Vx, *Vy = f->Vy, *Vz = f->Vz, *eta = f->eta;
    double* const* na = f->na;
    double* const* ns = f->ns;
    const double *wx = ib ? ns[OM_VCC] : NULL, *wy = ib ? ns[OM_CVC] : NULL, *wz = ib ? ns[OM_CCV] : NULL;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= nz + 1; ++k)
        for (int j = 1; j <= ny + 1; ++j)
            for (int i = 1; i <= nx + 1; ++i) {
                const long c = i + sy * j + sz * k;
                if (!ib) {
                    /* kernels_3d.jl:7-24 */
                    double exx = (Vx[c + 1] - Vx[c]) * ix;
                    double eyy = (Vy[c + sy] - Vy[c]) * iy;
                    double ezz = (Vz[c + sz] - Vz[c]) * iz;
                    double exy = 0.5 * ((Vy[c] - Vy[c - 1]) * ix + (Vx[c] - Vx[c - sy]) * iy);
                    double exz = 0.5 * ((Vz[c] - Vz[c - 1]) * ix + (Vx[c] - Vx[c - sz]) * iz);
                    double eyz = 0.5 * ((Vz[c] - Vz[c - sy]) * iy + (Vy[c] - Vy[c - sz]) * iz);
                    double divV = exx + eyy + ezz;
                    double ec = eta[c];
                    double exy_eta = 0.5 * (0.5 * (eta[c - 1 - sy] + eta[c - sy]) + 0.5 * (eta[c - 1] + eta[c]));
                    double exz_eta = 0.5 * (0.5 * (eta[c - 1 - sz] + eta[c - sz]) + 0.5 * (eta[c - 1] + eta[c]));
                    double eyz_eta = 0.5 * (0.5 * (eta[c - sy - sz] + eta[c - sz]) + 0.5 * (eta[c - sy] + eta[c]));
                    P[c] = P[c] - divV * ec * m->dPr;
                    txx[c] = txx[c] - (txx[c] - 2.0 * ec * (exx - divV / 3.0)) * m->dtau[0];
                    tyy[c] = tyy[c] - (tyy[c] - 2.0 * ec * (eyy - divV / 3.0)) * m->dtau[1];
                    tzz[c] = tzz[c] - (tzz[c] - 2.0 * ec * (ezz - divV / 3.0)) * m->dtau[2];
                    txy[c] = txy[c] - (txy[c] - 2.0 * exy_eta * exy) * m->dtau[3];
                    txz[c] = txz[c] - (txz[c] - 2.0 * exz_eta * exz) * m->dtau[4];
                    tyz[c] = tyz[c] - (tyz[c] - 2.0 * eyz_eta * eyz) * m->dtau[5];
                } else {
                    /* IB kernels_3d.jl:6-43 */
                    if (valid3(na, ns, L_CCC, c, sy, sz)) {
                        double exx = (wx[c + 1] * Vx[c + 1] - wx[c] * Vx[c]) * ix;
                        double eyy = (wy[c + sy] * Vy[c + sy] - wy[c] * Vy[c]) * iy;
                        double ezz = (wz[c + sz] * Vz[c + sz] - wz[c] * Vz[c]) * iz;
                        double divV = exx + eyy + ezz;
                        double ec = eta[c];
                        P[c] = P[c] - divV * ec * m->dPr;
                        txx[c] = txx[c] - (txx[c] - 2.0 * ec * (exx - divV / 3.0)) * m->dtau[0];
                        tyy[c] = tyy[c] - (tyy[c] - 2.0 * ec * (eyy - divV / 3.0)) * m->dtau[1];
                        tzz[c] = tzz[c] - (tzz[c] - 2.0 * ec * (ezz - divV / 3.0)) * m->dtau[2];
                    } else {
                        P[c] = 0.0; txx[c] = 0.0; tyy[c] = 0.0; tzz[c] = 0.0;
                    }
                    if (valid3(na, ns, L_VVC, c, sy, sz)) {
                        double exy = 0.5 * ((wy[c] * Vy[c] - wy[c - 1] * Vy[c - 1]) * ix +
                                            (wx[c] * Vx[c] - wx[c - sy] * Vx[c - sy]) * iy);
                        double e = 0.5 * (0.5 * (eta[c - 1 - sy] + eta[c - sy]) + 0.5 * (eta[c - 1] + eta[c]));
                        txy[c] = txy[c] - (txy[c] - 2.0 * e * exy) * m->dtau[3];
                    } else txy[c] = 0.0;
                    if (valid3(na, ns, L_VCV, c, sy, sz)) {
                        double exz = 0.5 * ((wz[c] * Vz[c] - wz[c - 1] * Vz[c - 1]) * ix +
                                            (wx[c] * Vx[c] - wx[c - sz] * Vx[c - sz]) * iz);
                        double e = 0.5 * (0.5 * (eta[c - 1 - sz] + eta[c - sz]) + 0.5 * (eta[c - 1] + eta[c]));
                        txz[c] = txz[c] - (txz[c] - 2.0 * e * exz) * m->dtau[4];
                    } else txz[c] = 0.0;
                    if (valid3(na, ns, L_CVV, c, sy, sz)) {
                        double eyz = 0.5 * ((wz[c] * Vz[c] - wz[c - sy] * Vz[c - sy]) * iy +
                                            (wy[c] * Vy[c] - wy[c - sz] * Vy[c - sz]) * iz);
                        double e = 0.5 * (0.5 * (eta[c - sy - sz] + eta[c - sz]) + 0.5 * (eta[c - sy] + eta[c]));
                        tyz[c] = tyz[c] - (tyz[c] - 2.0 * e * eyz) * m->dtau[5];
                    } else tyz[c] = 0.0;
                }
            }
}

static double tauII3(const or_fields* f, long c, long sy, long sz)
{
    /* kernels_3d.jl:35-38 */
    const double *txy = f->txy, *txz = f->txz, *tyz = f->tyz;
    double a = f->txx[c], b = f->tyy[c], d = f->tzz[c];
    double lxy = 0.5 * (0.5 * (txy[c] + txy[c + 1]) + 0.5 * (txy[c + sy] + txy[c + 1 + sy]));
    double lxz = 0.5 * (0.5 * (txz[c] + txz[c + 1]) + 0.5 * (txz[c + sz] + txz[c + 1 + sz]));
    double lyz = 0.5 * (0.5 * (tyz[c] + tyz[c + sy]) + 0.5 * (tyz[c + sz] + tyz[c + sy + sz]));
    return sqrt(0.5 * (a * a + b * b + d * d) + lxy * lxy + lxz * lxz + lyz * lyz);
}

static void velocity3(const or_model* m, or_fields* f, int ib)
{
    const int nx = m->n[0], ny = m->n[1], nz = m->n[2];
    const long sy = m->sy, sz = m->sz;
    const double ix = m->idx[0], iy = m->idx[1], iz = m->idx[2];
    const double *P = f->P, *txx = f->txx, *tyy = f->tyy, *tzz = f->tzz, *txy = f->txy, *txz = f->txz, *tyz = f->tyz;
    double *Vx = f->Vx, *Vy = f->Vy, *Vz = f->Vz, *eta_next = f->eta_next;
    const double* eta = f->eta;
    double* const* na = f->na;
    double* const* ns = f->ns;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= nz + 1; ++k)
        for (int j = 1; j <= ny + 1; ++j)
            for (int i = 1; i <= nx + 1; ++i) {
                const long c = i + sy * j + sz * k;
                double gx = m->grav_field ? f->gx[c] : m->grav[0];
                double gy = m->grav_field ? f->gy[c] : m->grav[1];
                double gz = m->grav_field ? f->gz[c] : m->grav[2];
                double ex = 0.5 * (eta[c - 1] + eta[c]);
                double ey = 0.5 * (eta[c - sy] + eta[c]);
                double ez = 0.5 * (eta[c - sz] + eta[c]);
                if (!ib) {
                    /* kernels_3d.jl:31-33 */
                    Vx[c] = Vx[c] + (-((P[c] - P[c - 1]) * ix) + (txx[c] - txx[c - 1]) * ix +
                                     (txy[c + sy] - txy[c]) * iy + (txz[c + sz] - txz[c]) * iz - gx) / ex * m->dV[0];
                    Vy[c] = Vy[c] + (-((P[c] - P[c - sy]) * iy) + (tyy[c] - tyy[c - sy]) * iy +
                                     (txy[c + 1] - txy[c]) * ix + (tyz[c + sz] - tyz[c]) * iz - gy) / ey * m->dV[1];
                    Vz[c] = Vz[c] + (-((P[c] - P[c - sz]) * iz) + (tzz[c] - tzz[c - sz]) * iz +
                                     (txz[c + 1] - txz[c]) * ix + (tyz[c + sy] - tyz[c]) * iy - gz) / ez * m->dV[2];
                } else {
                    /* IB kernels_3d.jl:50-67 */
                    const double *wc = na[OM_CCC], *wxy = na[OM_VVC], *wxz = na[OM_VCV], *wyz = na[OM_CVV];
                    if (valid3(ns, na, L_VCC, c, sy, sz))
                        Vx[c] = Vx[c] + (-((wc[c] * P[c] - wc[c - 1] * P[c - 1]) * ix) +
                                         (wc[c] * txx[c] - wc[c - 1] * txx[c - 1]) * ix +
                                         (wxy[c + sy] * txy[c + sy] - wxy[c] * txy[c]) * iy +
                                         (wxz[c + sz] * txz[c + sz] - wxz[c] * txz[c]) * iz - gx) /
                                            ex / ns[OM_VCC][c] * m->dV[0];
                    else Vx[c] = 0.0;
                    if (valid3(ns, na, L_CVC, c, sy, sz))
                        Vy[c] = Vy[c] + (-((wc[c] * P[c] - wc[c - sy] * P[c - sy]) * iy) +
                                         (wc[c] * tyy[c] - wc[c - sy] * tyy[c - sy]) * iy +
                                         (wxy[c + 1] * txy[c + 1] - wxy[c] * txy[c]) * ix +
                                         (wyz[c + sz] * tyz[c + sz] - wyz[c] * tyz[c]) * iz - gy) /
                                            ey / ns[OM_CVC][c] * m->dV[1];
                    else Vy[c] = 0.0;
                    if (valid3(ns, na, L_CCV, c, sy, sz))
                        Vz[c] = Vz[c] + (-((wc[c] * P[c] - wc[c - sz] * P[c - sz]) * iz) +
                                         (wc[c] * tzz[c] - wc[c - sz] * tzz[c - sz]) * iz +
                                         (wxz[c + 1] * txz[c + 1] - wxz[c] * txz[c]) * ix +
                                         (wyz[c + sy] * tyz[c + sy] - wyz[c] * tyz[c]) * iy - gz) /
                                            ez / ns[OM_CCV][c] * m->dV[2];
                    else Vz[c] = 0.0;
                }
                eta_next[c] = or_rheology(m, f, c, tauII3(f, c, sy, sz)); /* kernels_3d.jl:35-39, IB :69-73 */
            }
}

static void residuals3(const or_model* m, or_fields* f, int ib)
{
    const int nx = m->n[0], ny = m->n[1], nz = m->n[2];
    const long sy = m->sy, sz = m->sz;
    const double ix = m->idx[0], iy = m->idx[1], iz = m->idx[2];
    const double *P = f->P, *txx = f->txx, *tyy = f->tyy, *tzz = f->tzz, *txy = f->txy, *txz = f->txz, *tyz = f->tyz;
    const double *Vx = f->Vx, *Vy = f->Vy, *Vz = f->Vz;
    double *rP = f->rP, *rVx = f->rVx, *rVy = f->rVy, *rVz = f->rVz;
    double* const* na = f->na;
    double* const* ns = f->ns;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= nz + 1; ++k)
        for (int j = 1; j <= ny + 1; ++j)
            for (int i = 1; i <= nx + 1; ++i) {
                const long c = i + sy * j + sz * k;
                double gx = m->grav_field ? f->gx[c] : m->grav[0];
                double gy = m->grav_field ? f->gy[c] : m->grav[1];
                double gz = m->grav_field ? f->gz[c] : m->grav[2];
                if (!ib) {
                    /* kernels_3d.jl:70-74 */
                    rP[c] = (Vx[c + 1] - Vx[c]) * ix + (Vy[c + sy] - Vy[c]) * iy + (Vz[c + sz] - Vz[c]) * iz;
                    rVx[c] = -((P[c] - P[c - 1]) * ix) + (txx[c] - txx[c - 1]) * ix + (txy[c + sy] - txy[c]) * iy +
                             (txz[c + sz] - txz[c]) * iz - gx;
                    rVy[c] = -((P[c] - P[c - sy]) * iy) + (tyy[c] - tyy[c - sy]) * iy + (txy[c + 1] - txy[c]) * ix +
                             (tyz[c + sz] - tyz[c]) * iz - gy;
                    rVz[c] = -((P[c] - P[c - sz]) * iz) + (tzz[c] - tzz[c - sz]) * iz + (txz[c + 1] - txz[c]) * ix +
                             (tyz[c + sy] - tyz[c]) * iy - gz;
                } else {
                    /* IB kernels_3d.jl:122-142 */
                    const double *wx = ns[OM_VCC], *wy = ns[OM_CVC], *wz = ns[OM_CCV];
                    const double *wc = na[OM_CCC], *wxy = na[OM_VVC], *wxz = na[OM_VCV], *wyz = na[OM_CVV];
                    if (valid3(na, ns, L_CCC, c, sy, sz))
                        rP[c] = (wx[c + 1] * Vx[c + 1] - wx[c] * Vx[c]) * ix + (wy[c + sy] * Vy[c + sy] - wy[c] * Vy[c]) * iy +
                                (wz[c + sz] * Vz[c + sz] - wz[c] * Vz[c]) * iz;
                    else rP[c] = 0.0;
                    if (valid3(ns, na, L_VCC, c, sy, sz))
                        rVx[c] = -((wc[c] * P[c] - wc[c - 1] * P[c - 1]) * ix) + (wc[c] * txx[c] - wc[c - 1] * txx[c - 1]) * ix +
                                 (wxy[c + sy] * txy[c + sy] - wxy[c] * txy[c]) * iy +
                                 (wxz[c + sz] * txz[c + sz] - wxz[c] * txz[c]) * iz - gx;
                    else rVx[c] = 0.0;
                    if (valid3(ns, na, L_CVC, c, sy, sz))
                        rVy[c] = -((wc[c] * P[c] - wc[c - sy] * P[c - sy]) * iy) +
                                 (wc[c] * tyy[c] - wc[c - sy] * tyy[c - sy]) * iy +
                                 (wxy[c + 1] * txy[c + 1] - wxy[c] * txy[c]) * ix +
                                 (wyz[c + sz] * tyz[c + sz] - wyz[c] * tyz[c]) * iz - gy;
                    else rVy[c] = 0.0;
                    if (valid3(ns, na, L_CCV, c, sy, sz))
                        rVz[c] = -((wc[c] * P[c] - wc[c - sz] * P[c - sz]) * iz) +
                                 (wc[c] * tzz[c] - wc[c - sz] * tzz[c - sz]) * iz +
                                 (wxz[c + 1] * txz[c + 1] - wxz[c] * txz[c]) * ix +
                                 (wyz[c + sy] * tyz[c + sy] - wyz[c] * tyz[c]) * iy - gz;
                    else rVz[c] = 0.0;
                }
            }
}

static void tau3(const or_model* m, or_fields* f, int ib)
{
    /* compute_tau!: kernels_3d.jl:44-64 (IB: IB kernels_3d.jl:78-116, with the
     * undefined `Pr` location read as the Center location of tau.xx) */
    const int nx = m->n[0], ny = m->n[1], nz = m->n[2];
    const long sy = m->sy, sz = m->sz;
    const double ix = m->idx[0], iy = m->idx[1], iz = m->idx[2];
    double *txx = f->txx, *tyy = f->tyy, *tzz = f->tzz, *txy = f->txy, *txz = f->txz, *tyz = f->tyz;
    const double *Vx = f->Vx, *Vy = f->Vy, *Vz = f->Vz, *eta = f->eta;
    double* const* na = f->na;
    double* const* ns = f->ns;
    const double one = 1.0;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= nz + 1; ++k)
        for (int j = 1; j <= ny + 1; ++j)
            for (int i = 1; i <= nx + 1; ++i) {
                const long c = i + sy * j + sz * k;
                const double *wx = ib ? ns[OM_VCC] : NULL, *wy = ib ? ns[OM_CVC] : NULL, *wz = ib ? ns[OM_CCV] : NULL;
#define WX(q) (ib ? wx[q] : one)
#define WY(q) (ib ? wy[q] : one)
#define WZ(q) (ib ? wz[q] : one)
                double ec = eta[c];
                if (!ib || valid3(na, ns, L_CCC, c, sy, sz)) {
                    double exx = ib ? (WX(c + 1) * Vx[c + 1] - WX(c) * Vx[c]) * ix : (Vx[c + 1] - Vx[c]) * ix;
                    double eyy = ib ? (WY(c + sy) * Vy[c + sy] - WY(c) * Vy[c]) * iy : (Vy[c + sy] - Vy[c]) * iy;
                    double ezz = ib ? (WZ(c + sz) * Vz[c + sz] - WZ(c) * Vz[c]) * iz : (Vz[c + sz] - Vz[c]) * iz;
                    double divV = exx + eyy + ezz;
                    txx[c] = 2.0 * ec * (exx - divV / 3.0);
                    tyy[c] = 2.0 * ec * (eyy - divV / 3.0);
                    tzz[c] = 2.0 * ec * (ezz - divV / 3.0);
                } else {
                    txx[c] = 0.0; tyy[c] = 0.0; tzz[c] = 0.0;
                }
                if (!ib || valid3(na, ns, L_VVC, c, sy, sz)) {
                    double exy = ib ? 0.5 * ((WY(c) * Vy[c] - WY(c - 1) * Vy[c - 1]) * ix + (WX(c) * Vx[c] - WX(c - sy) * Vx[c - sy]) * iy)
                                    : 0.5 * ((Vy[c] - Vy[c - 1]) * ix + (Vx[c] - Vx[c - sy]) * iy);
                    double e = 0.5 * (0.5 * (eta[c - 1 - sy] + eta[c - sy]) + 0.5 * (eta[c - 1] + eta[c]));
                    txy[c] = 2.0 * e * exy;
                } else txy[c] = 0.0;
                if (!ib || valid3(na, ns, L_VCV, c, sy, sz)) {
                    double exz = ib ? 0.5 * ((WZ(c) * Vz[c] - WZ(c - 1) * Vz[c - 1]) * ix + (WX(c) * Vx[c] - WX(c - sz) * Vx[c - sz]) * iz)
                                    : 0.5 * ((Vz[c] - Vz[c - 1]) * ix + (Vx[c] - Vx[c - sz]) * iz);
                    double e = 0.5 * (0.5 * (eta[c - 1 - sz] + eta[c - sz]) + 0.5 * (eta[c - 1] + eta[c]));
                    txz[c] = 2.0 * e * exz;
                } else txz[c] = 0.0;
                if (!ib || valid3(na, ns, L_CVV, c, sy, sz)) {
                    double eyz = ib ? 0.5 * ((WZ(c) * Vz[c] - WZ(c - sy) * Vz[c - sy]) * iy + (WY(c) * Vy[c] - WY(c - sz) * Vy[c - sz]) * iz)
                                    : 0.5 * ((Vz[c] - Vz[c - sy]) * iy + (Vy[c] - Vy[c - sz]) * iz);
                    double e = 0.5 * (0.5 * (eta[c - sy - sz] + eta[c - sz]) + 0.5 * (eta[c - sy] + eta[c]));
                    tyz[c] = 2.0 * e * eyz;
                } else tyz[c] = 0.0;
#undef WX
#undef WY
#undef WZ
            }
}

/* ------------------------------------------------------------------ */
/* 2D kernels: kernels_2d.jl (FullStokes); the IB 2D path restates the  */
/* 3D IB kernels in two dimensions                                      */
/* ------------------------------------------------------------------ */
static void sigma2(const or_model* m, or_fields* f, int ib)
{
    const int nx = m->n[0], ny = m->n[1];
    const long sy = m->sy;
    const double ix = m->idx[0], iy = m->idx[1];
    double *P = f->P, *txx = f->txx, *tyy = f->tyy, *txy = f->txy;
    const double *Vx = f->Vx, *Vy = f->Vy, *eta = f->eta;
    double* const* na = f->na;
    double* const* ns = f->ns;
    const double *wx = ib ? ns[OM2_VC] : NULL, *wy = ib ? ns[OM2_CV] : NULL;
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny + 1; ++j)
        for (int i = 1; i <= nx + 1; ++i) {
            const long c = i + sy * j;
            double ec = eta[c];
            double exy_eta = 0.5 * (0.5 * (eta[c - 1 - sy] + eta[c - sy]) + 0.5 * (eta[c - 1] + eta[c]));
            if (!ib) {
                /* kernels_2d.jl:7-17 */
                double exx = (Vx[c + 1] - Vx[c]) * ix;
                double eyy = (Vy[c + sy] - Vy[c]) * iy;
                double exy = 0.5 * ((Vy[c] - Vy[c - 1]) * ix + (Vx[c] - Vx[c - sy]) * iy);
                double divV = exx + eyy;
                P[c] = P[c] - divV * ec * m->dPr;
                txx[c] = txx[c] - (txx[c] - 2.0 * ec * (exx - divV / 3.0)) * m->dtau[0];
                tyy[c] = tyy[c] - (tyy[c] - 2.0 * ec * (eyy - divV / 3.0)) * m->dtau[1];
                txy[c] = txy[c] - (txy[c] - 2.0 * exy_eta * exy) * m->dtau[2];
            } else {
                if (valid2(na, ns, L_CCC, c, sy)) {
                    double exx = (wx[c + 1] * Vx[c + 1] - wx[c] * Vx[c]) * ix;
                    double eyy = (wy[c + sy] * Vy[c + sy] - wy[c] * Vy[c]) * iy;
                    double divV = exx + eyy;
                    P[c] = P[c] - divV * ec * m->dPr;
                    txx[c] = txx[c] - (txx[c] - 2.0 * ec * (exx - divV / 3.0)) * m->dtau[0];
                    tyy[c] = tyy[c] - (tyy[c] - 2.0 * ec * (eyy - divV / 3.0)) * m->dtau[1];
                } else {
                    P[c] = 0.0; txx[c] = 0.0; tyy[c] = 0.0;
                }
                if (valid2(na, ns, L_VVC, c, sy)) {
                    double exy = 0.5 * ((wy[c] * Vy[c] - wy[c - 1] * Vy[c - 1]) * ix + (wx[c] * Vx[c] - wx[c - sy] * Vx[c - sy]) * iy);
                    txy[c] = txy[c] - (txy[c] - 2.0 * exy_eta * exy) * m->dtau[2];
                } else txy[c] = 0.0;
            }
        }
}

static double tauII2(const or_fields* f, long c, long sy)
{
    /* kernels_2d.jl:27 */
    const double* txy = f->txy;
    double a = f->txx[c], b = f->tyy[c];
    double l = 0.5 * (0.5 * (txy[c] + txy[c + 1]) + 0.5 * (txy[c + sy] + txy[c + 1 + sy]));
    return sqrt(0.5 * (a * a + b * b) + l * l);
}

static void velocity2(const or_model* m, or_fields* f, int ib)
{
    const int nx = m->n[0], ny = m->n[1];
    const long sy = m->sy;
    const double ix = m->idx[0], iy = m->idx[1];
    const double *P = f->P, *txx = f->txx, *tyy = f->tyy, *txy = f->txy, *eta = f->eta;
    double *Vx = f->Vx, *Vy = f->Vy, *eta_next = f->eta_next;
    double* const* na = f->na;
    double* const* ns = f->ns;
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny + 1; ++j)
        for (int i = 1; i <= nx + 1; ++i) {
            const long c = i + sy * j;
            double gx = m->grav_field ? f->gx[c] : m->grav[0];
            double gy = m->grav_field ? f->gy[c] : m->grav[1];
            /* kernels_2d.jl:24-25 interpolate eta to location(V.x) for both
             * components; each component here uses its own location */
            double ex = 0.5 * (eta[c - 1] + eta[c]);
            double ey = 0.5 * (eta[c - sy] + eta[c]);
            if (!ib) {
                Vx[c] = Vx[c] + (-((P[c] - P[c - 1]) * ix) + (txx[c] - txx[c - 1]) * ix + (txy[c + sy] - txy[c]) * iy - gx) / ex * m->dV[0];
                Vy[c] = Vy[c] + (-((P[c] - P[c - sy]) * iy) + (tyy[c] - tyy[c - sy]) * iy + (txy[c + 1] - txy[c]) * ix - gy) / ey * m->dV[1];
            } else {
                const double *wc = na[OM2_CC], *wv = na[OM2_VV];
                if (valid2(ns, na, L_VCC, c, sy))
                    Vx[c] = Vx[c] + (-((wc[c] * P[c] - wc[c - 1] * P[c - 1]) * ix) + (wc[c] * txx[c] - wc[c - 1] * txx[c - 1]) * ix +
                                     (wv[c + sy] * txy[c + sy] - wv[c] * txy[c]) * iy - gx) / ex / ns[OM2_VC][c] * m->dV[0];
                else Vx[c] = 0.0;
                if (valid2(ns, na, L_CVC, c, sy))
                    Vy[c] = Vy[c] + (-((wc[c] * P[c] - wc[c - sy] * P[c - sy]) * iy) + (wc[c] * tyy[c] - wc[c - sy] * tyy[c - sy]) * iy +
                                     (wv[c + 1] * txy[c + 1] - wv[c] * txy[c]) * ix - gy) / ey / ns[OM2_CV][c] * m->dV[1];
                else Vy[c] = 0.0;
            }
            eta_next[c] = or_rheology(m, f, c, tauII2(f, c, sy));
        }
}

static void residuals2(const or_model* m, or_fields* f, int ib)
{
    const int nx = m->n[0], ny = m->n[1];
    const long sy = m->sy;
    const double ix = m->idx[0], iy = m->idx[1];
    const double *P = f->P, *txx = f->txx, *tyy = f->tyy, *txy = f->txy, *Vx = f->Vx, *Vy = f->Vy;
    double *rP = f->rP, *rVx = f->rVx, *rVy = f->rVy;
    double* const* na = f->na;
    double* const* ns = f->ns;
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny + 1; ++j)
        for (int i = 1; i <= nx + 1; ++i) {
            const long c = i + sy * j;
            double gx = m->grav_field ? f->gx[c] : m->grav[0];
            double gy = m->grav_field ? f->gy[c] : m->grav[1];
            if (!ib) {
                /* kernels_2d.jl:50-53 */
                rP[c] = (Vx[c + 1] - Vx[c]) * ix + (Vy[c + sy] - Vy[c]) * iy;
                rVx[c] = -((P[c] - P[c - 1]) * ix) + (txx[c] - txx[c - 1]) * ix + (txy[c + sy] - txy[c]) * iy - gx;
                rVy[c] = -((P[c] - P[c - sy]) * iy) + (tyy[c] - tyy[c - sy]) * iy + (txy[c + 1] - txy[c]) * ix - gy;
            } else {
                const double *wx = ns[OM2_VC], *wy = ns[OM2_CV], *wc = na[OM2_CC], *wv = na[OM2_VV];
                if (valid2(na, ns, L_CCC, c, sy))
                    rP[c] = (wx[c + 1] * Vx[c + 1] - wx[c] * Vx[c]) * ix + (wy[c + sy] * Vy[c + sy] - wy[c] * Vy[c]) * iy;
                else rP[c] = 0.0;
                if (valid2(ns, na, L_VCC, c, sy))
                    rVx[c] = -((wc[c] * P[c] - wc[c - 1] * P[c - 1]) * ix) + (wc[c] * txx[c] - wc[c - 1] * txx[c - 1]) * ix +
                             (wv[c + sy] * txy[c + sy] - wv[c] * txy[c]) * iy - gx;
                else rVx[c] = 0.0;
                if (valid2(ns, na, L_CVC, c, sy))
                    rVy[c] = -((wc[c] * P[c] - wc[c - sy] * P[c - sy]) * iy) + (wc[c] * tyy[c] - wc[c - sy] * tyy[c - sy]) * iy +
                             (wv[c + 1] * txy[c + 1] - wv[c] * txy[c]) * ix - gy;
                else rVy[c] = 0.0;
            }
        }
}

static void tau2(const or_model* m, or_fields* f, int ib)
{
    /* compute_tau!: kernels_2d.jl:33-44 */
    const int nx = m->n[0], ny = m->n[1];
    const long sy = m->sy;
    const double ix = m->idx[0], iy = m->idx[1];
    double *txx = f->txx, *tyy = f->tyy, *txy = f->txy;
    const double *Vx = f->Vx, *Vy = f->Vy, *eta = f->eta;
    double* const* na = f->na;
    double* const* ns = f->ns;
    for (int j = 1; j <= ny + 1; ++j)
        for (int i = 1; i <= nx + 1; ++i) {
            const long c = i + sy * j;
            double ec = eta[c];
            double exy_eta = 0.5 * (0.5 * (eta[c - 1 - sy] + eta[c - sy]) + 0.5 * (eta[c - 1] + eta[c]));
            double wxr = ib ? ns[OM2_VC][c + 1] : 1.0, wxc = ib ? ns[OM2_VC][c] : 1.0, wxs = ib ? ns[OM2_VC][c - sy] : 1.0;
            double wyu = ib ? ns[OM2_CV][c + sy] : 1.0, wyc = ib ? ns[OM2_CV][c] : 1.0, wyw = ib ? ns[OM2_CV][c - 1] : 1.0;
            if (!ib || valid2(na, ns, L_CCC, c, sy)) {
                double exx = ib ? (wxr * Vx[c + 1] - wxc * Vx[c]) * ix : (Vx[c + 1] - Vx[c]) * ix;
                double eyy = ib ? (wyu * Vy[c + sy] - wyc * Vy[c]) * iy : (Vy[c + sy] - Vy[c]) * iy;
                double divV = exx + eyy;
                txx[c] = 2.0 * ec * (exx - divV / 3.0);
                tyy[c] = 2.0 * ec * (eyy - divV / 3.0);
            } else {
                txx[c] = 0.0; tyy[c] = 0.0;
            }
            if (!ib || valid2(na, ns, L_VVC, c, sy)) {
                double exy = ib ? 0.5 * ((wyc * Vy[c] - wyw * Vy[c - 1]) * ix + (wxc * Vx[c] - wxs * Vx[c - sy]) * iy)
                                : 0.5 * ((Vy[c] - Vy[c - 1]) * ix + (Vx[c] - Vx[c - sy]) * iy);
                txy[c] = 2.0 * exy_eta * exy;
            } else txy[c] = 0.0;
        }
}

/* ------------------------------------------------------------------ */
/* model-level entry points: Isothermal.jl:91-125                       */
/* ------------------------------------------------------------------ */
void or_update_sigma(const or_model* m, or_fields* f, int ib)
{
    if (m->ndims == 3) sigma3(m, f, ib); else sigma2(m, f, ib);
}

void or_update_V(const or_model* m, or_fields* f, int ib)
{
    if (m->ndims == 3) velocity3(m, f, ib); else velocity2(m, f, ib);
}

void or_advance(const or_model* m, or_fields* f, int niter, int ib)
{
    for (int it = 0; it < niter; ++it) {
        or_update_sigma(m, f, ib);
        or_bc_stress(m, f);
        or_update_V(m, f, ib);
        or_bc_velocity(m, f);
        double* t = f->eta; /* Isothermal.jl:110 */
        f->eta = f->eta_next;
        f->eta_next = t;
    }
}

void or_compute_residuals(const or_model* m, or_fields* f, int ib)
{
    if (m->ndims == 3) residuals3(m, f, ib); else residuals2(m, f, ib);
    or_bc_residual(m, f);
}

void or_compute_tau(const or_model* m, or_fields* f, int ib)
{
    if (m->ndims == 3) tau3(m, f, ib); else tau2(m, f, ib);
}

static double nanmax_abs(double acc, double v)
{
    double a = fabs(v);
    if (a != a || acc != acc) return NAN;
    return a > acc ? a : acc;
}

static double norm_interior(const or_model* m, const double* a, int L)
{
    int hi[3] = {1, 1, 1};
    for (int d = 0; d < m->ndims; ++d) hi[d] = m->n[d] + ((L >> d) & 1);
    int klo = (m->ndims == 3) ? 1 : 0, khi = (m->ndims == 3) ? hi[2] : 0;
    double acc = 0.0;
    for (int k = klo; k <= khi; ++k)
        for (int j = 1; j <= hi[1]; ++j)
            for (int i = 1; i <= hi[0]; ++i) acc = nanmax_abs(acc, a[i + m->sy * j + m->sz * k]);
    return acc;
}

void or_residual_norms(const or_model* m, const or_fields* f, double out[4])
{
    out[0] = norm_interior(m, f->rP, L_CCC);
    out[1] = norm_interior(m, f->rVx, L_VCC);
    out[2] = norm_interior(m, f->rVy, L_CVC);
    out[3] = (m->ndims == 3) ? norm_interior(m, f->rVz, L_CCV) : 0.0;
}

int or_solve(const or_model* m, or_fields* f, int max_iter, int ncheck, double tol, int ib, double norms[4])
{
    /* driver loop: tm_stokes_chmy.jl:96-108 with a tolerance exit */
    for (int it = 1; it <= max_iter; ++it) {
        or_advance(m, f, 1, ib);
        if (it % ncheck == 0) {
            or_compute_residuals(m, f, ib);
            or_residual_norms(m, f, norms);
            int conv = 1;
            for (int q = 0; q < 4; ++q) {
                if (!isfinite(norms[q])) return -1;
                if (!(norms[q] < tol)) conv = 0;
            }
            if (conv) return it;
        }
    }
    return max_iter;
}

/* ------------------------------------------------------------------ */
/* volume fractions: volume_fractions.jl:1-127                          */
/* ------------------------------------------------------------------ */
static double perturb(double p)
{
    return fabs(p) > 1e-20 ? p : (p > 0.0 ? 1e-20 : -1e-20);
}

typedef struct { double x, y, z; } v3;

static v3 vsub(v3 a, v3 b) { v3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }

/* det of the 3x3 matrix with columns a, b, c, by cofactors of the first row */
static double det3(v3 a, v3 b, v3 c)
{
    return a.x * (b.y * c.z - c.y * b.z) - b.x * (a.y * c.z - c.y * a.z) + c.x * (a.y * b.z - b.y * a.z);
}

static double tetvol(v3 a, v3 b, v3 c, v3 d)
{
    return (1.0 / 6.0) * fabs(det3(vsub(b, a), vsub(c, a), vsub(d, a))); /* volume_fractions.jl:5 */
}

static v3 vij3(const double* p, const v3* v, int i, int j)
{
    /* volume_fractions.jl:49 */
    double wi = p[i] / (p[i] - p[j]);
    double wj = p[j] / (p[i] - p[j]);
    v3 r = {v[j].x * wi - v[i].x * wj, v[j].y * wi - v[i].y * wj, v[j].z * wi - v[i].z * wj};
    return r;
}

static double tetcut(const double p[4], const v3 v[4])
{
    /* volume_fractions.jl:48-103; indices 0-based */
    int nneg = (p[0] < 0) + (p[1] < 0) + (p[2] < 0) + (p[3] < 0);
#define VIJ(a, b) vij3(p, v, a, b)
    if (nneg == 0) return 0.0;
    if (nneg == 1) {
        if (p[0] < 0) return tetvol(v[0], VIJ(0, 1), VIJ(0, 2), VIJ(0, 3));
        if (p[1] < 0) return tetvol(v[1], VIJ(1, 0), VIJ(1, 2), VIJ(1, 3));
        if (p[2] < 0) return tetvol(v[2], VIJ(2, 0), VIJ(2, 1), VIJ(2, 3));
        return tetvol(v[3], VIJ(3, 0), VIJ(3, 1), VIJ(3, 2));
    }
    if (nneg == 2) {
        if (p[0] < 0 && p[1] < 0)
            return tetvol(v[0], v[1], VIJ(0, 2), VIJ(1, 3)) + tetvol(VIJ(1, 2), v[1], VIJ(0, 2), VIJ(1, 3)) +
                   tetvol(v[0], VIJ(0, 3), VIJ(0, 2), VIJ(1, 3));
        if (p[0] < 0 && p[2] < 0)
            return tetvol(v[0], v[2], VIJ(0, 3), VIJ(2, 1)) + tetvol(VIJ(2, 3), v[2], VIJ(0, 3), VIJ(2, 1)) +
                   tetvol(v[0], VIJ(0, 1), VIJ(0, 3), VIJ(2, 1));
        if (p[0] < 0 && p[3] < 0)
            return tetvol(v[0], v[3], VIJ(0, 1), VIJ(3, 2)) + tetvol(VIJ(3, 1), v[3], VIJ(0, 1), VIJ(3, 2)) +
                   tetvol(v[0], VIJ(0, 2), VIJ(0, 1), VIJ(3, 2));
        if (p[1] < 0 && p[2] < 0)
            return tetvol(v[2], v[1], VIJ(2, 0), VIJ(1, 3)) + tetvol(VIJ(1, 0), v[1], VIJ(2, 0), VIJ(1, 3)) +
                   tetvol(v[2], VIJ(2, 3), VIJ(2, 0), VIJ(1, 3));
        if (p[1] < 0 && p[3] < 0)
            return tetvol(v[3], v[1], VIJ(3, 0), VIJ(1, 2)) + tetvol(VIJ(1, 0), v[1], VIJ(3, 0), VIJ(1, 2)) +
                   tetvol(v[3], VIJ(3, 2), VIJ(3, 0), VIJ(1, 2));
        return tetvol(v[2], v[3], VIJ(2, 0), VIJ(3, 1)) + tetvol(VIJ(3, 0), v[3], VIJ(2, 0), VIJ(3, 1)) +
               tetvol(v[2], VIJ(2, 1), VIJ(2, 0), VIJ(3, 1));
    }
    if (nneg == 3) {
        double tot = tetvol(v[0], v[1], v[2], v[3]);
        if (p[0] >= 0) return tot - tetvol(v[0], VIJ(0, 1), VIJ(0, 2), VIJ(0, 3));
        if (p[1] >= 0) return tot - tetvol(v[1], VIJ(1, 0), VIJ(1, 2), VIJ(1, 3));
        if (p[2] >= 0) return tot - tetvol(v[2], VIJ(2, 0), VIJ(2, 1), VIJ(2, 3));
        return tot - tetvol(v[3], VIJ(3, 0), VIJ(3, 1), VIJ(3, 2));
    }
    return tetvol(v[0], v[1], v[2], v[3]);
#undef VIJ
}

double or_psi2omega3(const double psi8[8], const double d[3])
{
    /* volume_fractions.jl:106-115,124-127; psi8 = (000,100,010,110,001,101,011,111) in (x,y,z) */
    const v3 v000 = {0, 0, 0}, v001 = {d[0], 0, 0}, v010 = {0, d[1], 0}, v011 = {d[0], d[1], 0};
    const v3 v100 = {0, 0, d[2]}, v101 = {d[0], 0, d[2]}, v110 = {0, d[1], d[2]}, v111 = {d[0], d[1], d[2]};
    double p[8];
    for (int q = 0; q < 8; ++q) p[q] = perturb(psi8[q]);
    double t1[4] = {p[0], p[4], p[2], p[1]};
    v3 w1[4] = {v000, v100, v010, v001};
    double t2[4] = {p[6], p[4], p[2], p[7]};
    v3 w2[4] = {v110, v100, v010, v111};
    double t3[4] = {p[5], p[4], p[7], p[1]};
    v3 w3[4] = {v101, v100, v111, v001};
    double t4[4] = {p[3], p[7], p[2], p[1]};
    v3 w4[4] = {v011, v111, v010, v001};
    double t5[4] = {p[7], p[4], p[2], p[1]};
    v3 w5[4] = {v111, v100, v010, v001};
    double vol = tetcut(t1, w1) + tetcut(t2, w2) + tetcut(t3, w3) + tetcut(t4, w4) + tetcut(t5, w5);
    return vol / (d[0] * d[1] * d[2]);
}

typedef struct { double x, y; } v2;

static double trivol(v2 a, v2 b, v2 c)
{
    /* volume_fractions.jl:4: 0.5*|det([c-a b-a])| */
    double ux = c.x - a.x, uy = c.y - a.y, wx = b.x - a.x, wy = b.y - a.y;
    return 0.5 * fabs(ux * wy - wx * uy);
}

static v2 vij2(const double* p, const v2* v, int i, int j)
{
    double wi = p[i] / (p[i] - p[j]);
    double wj = p[j] / (p[i] - p[j]);
    v2 r = {v[j].x * wi - v[i].x * wj, v[j].y * wi - v[i].y * wj};
    return r;
}

static double tricut(const double p[3], const v2 v[3])
{
    /* volume_fractions.jl:8-37 */
    if (p[0] < 0 && p[1] < 0 && p[2] < 0) return trivol(v[0], v[1], v[2]);
    if (p[0] > 0 && p[1] > 0 && p[2] > 0) return 0.0;
    v2 v12 = vij2(p, v, 0, 1), v13 = vij2(p, v, 0, 2), v23 = vij2(p, v, 1, 2);
    if (p[0] < 0) {
        if (p[1] < 0) return trivol(v[0], v23, v13) + trivol(v[0], v[1], v23);
        if (p[2] < 0) return trivol(v[2], v12, v23) + trivol(v[2], v[0], v12);
        return trivol(v[0], v12, v13);
    }
    if (p[1] < 0) {
        if (p[2] < 0) return trivol(v[1], v13, v12) + trivol(v[1], v[2], v13);
        return trivol(v12, v[1], v23);
    }
    return trivol(v13, v23, v[2]);
}

double or_psi2omega2(const double psi4[4], const double d[2])
{
    /* volume_fractions.jl:40-45,118-121; psi4 = (00, 10, 01, 11) in (x,y) */
    const v2 v00 = {0, 0}, v01 = {d[0], 0}, v10 = {0, d[1]}, v11 = {d[0], d[1]};
    double p00 = perturb(psi4[0]), p01 = perturb(psi4[1]), p10 = perturb(psi4[2]), p11 = perturb(psi4[3]);
    double a[3] = {p00, p01, p11};
    v2 va[3] = {v00, v01, v11};
    double b[3] = {p00, p11, p10};
    v2 vb[3] = {v00, v11, v10};
    return (tricut(a, va) + tricut(b, vb)) / (d[0] * d[1]);
}

/* psi lerps to dual locations; 'mask' bit d set = average along dim d */
static double psi_lerp3(const double* psi, long c, int mask, long sy, long sz)
{
    /* gather the 2x2x2 (or smaller) stencil, x innermost */
    int nxp = (mask & 1) ? 2 : 1, nyp = (mask & 2) ? 2 : 1, nzp = (mask & 4) ? 2 : 1;
    double zs[2];
    for (int kk = 0; kk < nzp; ++kk) {
        double ys[2];
        for (int jj = 0; jj < nyp; ++jj) {
            long b = c + jj * sy + kk * sz;
            ys[jj] = (nxp == 2) ? 0.5 * (psi[b] + psi[b + 1]) : psi[b];
        }
        zs[kk] = (nyp == 2) ? 0.5 * (ys[0] + ys[1]) : ys[0];
    }
    return (nzp == 2) ? 0.5 * (zs[0] + zs[1]) : zs[0];
}

static double psi_lerp2(const double* psi, long c, int mask, long sy)
{
    int nxp = (mask & 1) ? 2 : 1, nyp = (mask & 2) ? 2 : 1;
    double ys[2];
    for (int jj = 0; jj < nyp; ++jj) {
        long b = c + jj * sy;
        ys[jj] = (nxp == 2) ? 0.5 * (psi[b] + psi[b + 1]) : psi[b];
    }
    return (nyp == 2) ? 0.5 * (ys[0] + ys[1]) : ys[0];
}

void or_omega_from_psi(const or_model* m, const double* psi, double* om[8])
{
    /* volume_fractions_kernels.jl:2-28 (2D), 31-97 (3D) */
    const long sy = m->sy, sz = m->sz;
    if (m->ndims == 3) {
        /* per component: lerp mask of psi and the corner offset (-1 or 0) per dim */
        const int lmask[8] = {0, 7, 1, 2, 4, 3, 5, 6};
        const int off[8][3] = {{0, 0, 0}, {-1, -1, -1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {-1, -1, 0}, {-1, 0, -1}, {0, -1, -1}};
        const int nx = m->n[0], ny = m->n[1], nz = m->n[2];
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 1; k <= nz + 1; ++k)
            for (int j = 1; j <= ny + 1; ++j)
                for (int i = 1; i <= nx + 1; ++i) {
                    const long c = i + sy * j + sz * k;
                    for (int q = 0; q < 8; ++q) {
                        double p8[8];
                        for (int cz = 0; cz < 2; ++cz)
                            for (int cy = 0; cy < 2; ++cy)
                                for (int cx = 0; cx < 2; ++cx) {
                                    long b = c + (off[q][0] + cx) + sy * (off[q][1] + cy) + sz * (off[q][2] + cz);
                                    p8[cx + 2 * cy + 4 * cz] = psi_lerp3(psi, b, lmask[q], sy, sz);
                                }
                        om[q][c] = or_psi2omega3(p8, m->dx);
                    }
                }
    } else {
        const int lmask[4] = {0, 3, 1, 2};
        const int off[4][2] = {{0, 0}, {-1, -1}, {-1, 0}, {0, -1}};
        const int nx = m->n[0], ny = m->n[1];
        for (int j = 1; j <= ny + 1; ++j)
            for (int i = 1; i <= nx + 1; ++i) {
                const long c = i + sy * j;
                for (int q = 0; q < 4; ++q) {
                    double p4[4];
                    for (int cy = 0; cy < 2; ++cy)
                        for (int cx = 0; cx < 2; ++cx) {
                            long b = c + (off[q][0] + cx) + sy * (off[q][1] + cy);
                            p4[cx + 2 * cy] = psi_lerp2(psi, b, lmask[q], sy);
                        }
                    om[q][c] = or_psi2omega2(p4, m->dx);
                }
            }
    }
}

/* ------------------------------------------------------------------ */
/* signed distance to a DEM: signed_distances.jl:22-92,                 */
/* compute_levelsets.jl:1-35                                            */
/* ------------------------------------------------------------------ */
static double dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static v3 cross3(v3 a, v3 b)
{
    v3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    return r;
}
static v3 vscale(v3 a, double s) { v3 r = {a.x * s, a.y * s, a.z * s}; return r; }
static double sgn(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }
static double clamp01(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }

static double ud_triangle(v3 p, v3 a, v3 b, v3 c)
{
    /* signed_distances.jl:25-43 */
    v3 ba = vsub(b, a), pa = vsub(p, a), cb = vsub(c, b), pb = vsub(p, b), ac = vsub(a, c), pc = vsub(p, c);
    v3 nor = cross3(ba, ac);
    double s = sgn(dot3(cross3(ba, nor), pa)) + sgn(dot3(cross3(cb, nor), pb)) + sgn(dot3(cross3(ac, nor), pc));
    if (s < 2.0) {
        v3 e1 = vsub(vscale(ba, clamp01(dot3(ba, pa) / dot3(ba, ba))), pa);
        v3 e2 = vsub(vscale(cb, clamp01(dot3(cb, pb) / dot3(cb, cb))), pb);
        v3 e3 = vsub(vscale(ac, clamp01(dot3(ac, pc) / dot3(ac, ac))), pc);
        double d1 = dot3(e1, e1), d2 = dot3(e2, e2), d3 = dot3(e3, e3);
        double mn = d1 < d2 ? d1 : d2;
        mn = mn < d3 ? mn : d3;
        return sqrt(mn);
    }
    double np = dot3(nor, pa);
    return sqrt(np * np / dot3(nor, nor));
}

static double sign_triangle(v3 p, v3 a, v3 b, v3 c)
{
    double x = dot3(vsub(p, a), cross3(vsub(b, a), vsub(c, a)));
    return x == 0.0 ? 1.0 : sgn(x); /* signed_distances.jl:22-23 */
}

typedef struct {
    const double* dem;
    int nv[2];       /* vertex counts dnx+1, dny+1 */
    long sy;         /* dem row stride dnx+3 */
    double o[2], d[2];
} dem_grid;

static v3 sample_dem(const dem_grid* g, int i, int j)
{
    v3 r = {g->o[0] + (double)(i - 1) * g->d[0], g->o[1] + (double)(j - 1) * g->d[1], g->dem[i + g->sy * j]};
    return r;
}

static int closest_vertex(const dem_grid* g, double p, int d)
{
    /* signed_distances.jl:45-53 */
    int I = (int)floor((p - g->o[d]) / g->d[d]) + 1;
    if (I < 1) I = 1;
    if (I > g->nv[d]) I = g->nv[d];
    return I;
}

static void pair_dist(const dem_grid* g, v3 p, int i, int j, double* ud, double* sg)
{
    /* signed_distances.jl:58-72 */
    v3 a = sample_dem(g, i, j), b = sample_dem(g, i + 1, j), c = sample_dem(g, i, j + 1), e = sample_dem(g, i + 1, j + 1);
    double u1 = ud_triangle(p, a, b, c), u2 = ud_triangle(p, c, b, e);
    *ud = u1 < u2 ? u1 : u2;
    if (sg) *sg = sign_triangle(p, a, b, c);
}

void or_levelset_from_dem(const or_model* m, double* psi, const double* dem, const int dem_n[2],
                          const double dem_origin[2], const double dem_extent[2], const double R[9])
{
    dem_grid g;
    g.dem = dem;
    g.nv[0] = dem_n[0] + 1;
    g.nv[1] = dem_n[1] + 1;
    g.sy = dem_n[0] + 3;
    for (int d = 0; d < 2; ++d) {
        g.o[d] = dem_origin[d];
        g.d[d] = dem_extent[d] / (double)dem_n[d];
    }
    double dmax = m->dx[0];
    for (int d = 1; d < 3; ++d) dmax = m->dx[d] > dmax ? m->dx[d] : dmax;
    const double cutoff = 4.0 * dmax; /* compute_levelsets.jl:22 */
    const long sy = m->sy, sz = m->sz;
    const int nx = m->n[0], ny = m->n[1], nz = m->n[2];
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
    for (int k = 1; k <= nz + 1; ++k)
        for (int j = 1; j <= ny + 1; ++j)
            for (int i = 1; i <= nx + 1; ++i) {
                double x = m->origin[0] + (double)(i - 1) * m->dx[0];
                double y = m->origin[1] + (double)(j - 1) * m->dx[1];
                double z = m->origin[2] + (double)(k - 1) * m->dx[2];
                v3 p = {x, y, z};
                if (R) {
                    p.x = R[0] * x + R[1] * y + R[2] * z;
                    p.y = R[3] * x + R[4] * y + R[5] * z;
                    p.z = R[6] * x + R[7] * y + R[8] * z;
                }
                /* sd_dem: signed_distances.jl:78-92 */
                double px = p.x, py = p.y;
                double lo0 = g.o[0], hi0 = g.o[0] + dem_extent[0], lo1 = g.o[1], hi1 = g.o[1] + dem_extent[1];
                px = px < lo0 ? lo0 : (px > hi0 ? hi0 : px);
                py = py < lo1 ? lo1 : (py > hi1 ? hi1 : py);
                int bl0 = closest_vertex(&g, px - cutoff, 0), bl1 = closest_vertex(&g, py - cutoff, 1);
                int tr0 = closest_vertex(&g, px + cutoff, 0), tr1 = closest_vertex(&g, py + cutoff, 1);
                int ic0 = closest_vertex(&g, px, 0), ic1 = closest_vertex(&g, py, 1);
                double ud, sg;
                pair_dist(&g, p, ic0, ic1, &ud, &sg);
                for (int jv = bl1; jv <= tr1; ++jv)
                    for (int iv = bl0; iv <= tr0; ++iv) {
                        if (iv == ic0 && jv == ic1) continue;
                        double u;
                        pair_dist(&g, p, iv, jv, &u, NULL);
                        ud = u < ud ? u : ud;
                    }
                psi[i + sy * j + sz * k] = ud * sg;
            }
    or_bc_neumann_vertex(m, psi);
}

/* sd_poly: signed distance of p to a closed polygon, negative inside
 * (even-odd rule), signed_distances.jl:3-19 */
double or_sd_poly(double px, double py, const double* xy, int npts)
{
    double dx0 = px - xy[0], dy0 = py - xy[1];
    double d = dx0 * dx0 + dy0 * dy0;
    double s = 1.0;
    int j = npts - 1;
    for (int i = 0; i < npts; ++i) {
        const double ex = xy[2 * j] - xy[2 * i], ey = xy[2 * j + 1] - xy[2 * i + 1];
        const double wx = px - xy[2 * i], wy = py - xy[2 * i + 1];
        double t = (wx * ex + wy * ey) / (ex * ex + ey * ey);
        t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
        const double bx = wx - ex * t, by = wy - ey * t;
        const double bb = bx * bx + by * by;
        d = bb < d ? bb : d;
        const int c1 = py >= xy[2 * i + 1], c2 = py < xy[2 * j + 1], c3 = ex * wy > ey * wx;
        if ((c1 && c2 && c3) || (!c1 && !c2 && !c3)) s = -s;
        j = i;
    }
    return s * sqrt(d);
}

/* 2D level set of a polygon at every vertex, then Neumann on psi (the 2D
 * counterpart of compute_levelsets.jl:16-25 with sd_poly as the distance) */
void or_levelset_from_polygon(const or_model* m, double* psi, const double* xy, int npts)
{
    const long sy = m->sy;
    for (int j = 1; j <= m->n[1] + 1; ++j)
        for (int i = 1; i <= m->n[0] + 1; ++i) {
            const double x = m->origin[0] + (double)(i - 1) * m->dx[0];
            const double y = m->origin[1] + (double)(j - 1) * m->dx[1];
            psi[i + sy * j] = or_sd_poly(x, y, xy, npts);
        }
    or_bc_neumann_vertex(m, psi);
}

void or_invert_levelset(const or_model* m, double* psi)
{
    /* compute_levelsets.jl:10-14,32-35 */
    const long sy = m->sy, sz = m->sz;
    const int nz = (m->ndims == 3) ? m->n[2] : 0;
    for (int k = (m->ndims == 3) ? 1 : 0; k <= ((m->ndims == 3) ? nz + 1 : 0); ++k)
        for (int j = 1; j <= m->n[1] + 1; ++j)
            for (int i = 1; i <= m->n[0] + 1; ++i) psi[i + sy * j + sz * k] = -psi[i + sy * j + sz * k];
    or_bc_neumann_vertex(m, psi);
}

/* ------------------------------------------------------------------------
 * PT heat diffusion (src/Models/Diffusion/Heat/kernels.jl:3-29, with the
 * Value / Flux boundary conditions of Heat/boundary_conditions.jl:1-69:
 * Value -> DirichletBC{HalfCell} on T, Flux -> DirichletBC{FullCell} on q)
 * ---------------------------------------------------------------------- */
static long heat_stride(const or_heat* h, int d) { return d == 0 ? 1 : (d == 1 ? h->sy : h->sz); }

/* q.x[I] -= (q.x[I] + lam * dvx(T, I) / dx) * dtq  (kernels.jl:5-9) */
static double heat_flux_new(const or_heat* h, const or_heat_fields* f, int d, long c)
{
    const double dT = h->lam * (f->T[c] - f->T[c - heat_stride(h, d)]) / h->dx[d];
    return f->q[d][c] - (f->q[d][c] + dT) * h->dtq;
}

static long heat_index(const or_heat* h, int i, int j, int k) { return i + h->sy * j + h->sz * k; }

void or_heat_bc(const or_heat* h, or_heat_fields* f)
{
    /* x, then y, then z; every tangential index of the padded box */
    const int nd = h->ndims;
    for (int d = 0; d < nd; ++d) {
        const int e1 = d == 0 ? 1 : 0, e2 = nd == 3 ? (d == 2 ? 1 : 2) : -1;
        const int m1 = h->n[e1] + 2, m2 = e2 >= 0 ? h->n[e2] + 2 : 1;
        const long st = heat_stride(h, d), s1 = heat_stride(h, e1), s2 = e2 >= 0 ? heat_stride(h, e2) : 0;
        for (int s = 0; s < 2; ++s) {
            if (h->bc_kind[d][s] != OR_HEAT_VALUE) continue;
            const double v = h->bc_val[d][s];
            const long g = s == 0 ? 0 : h->n[d] + 1, src = s == 0 ? 1 : h->n[d];
            for (int b = 0; b < m2; ++b)
                for (int a = 0; a < m1; ++a) {
                    const long base = a * s1 + b * s2;
                    f->T[base + g * st] = 2.0 * v - f->T[base + src * st];
                }
        }
    }
}

void or_heat_advance(const or_heat* h, or_heat_fields* f, double dt, int niter)
{
    const int nd = h->ndims;
    const int nz = nd == 3 ? h->n[2] : 1, kz0 = nd == 3 ? 1 : 0;
    const long total = h->sz ? h->sz * (h->n[2] + 3) : h->sy * (h->n[1] + 3);
    double* qn[3] = {NULL, NULL, NULL};
    for (int d = 0; d < nd; ++d) qn[d] = (double*)malloc(sizeof(double) * total);
    for (int it = 0; it < niter; ++it) {
        /* update_q! over every face of each component, then the flux BCs */
        for (int d = 0; d < nd; ++d) {
            memcpy(qn[d], f->q[d], sizeof(double) * total);
            const int hx = h->n[0] + (d == 0), hy = h->n[1] + (d == 1), hz = nd == 3 ? h->n[2] + (d == 2) : 1;
            for (int k = kz0; k < kz0 + hz; ++k)
                for (int j = 1; j <= hy; ++j)
                    for (int i = 1; i <= hx; ++i) {
                        const long c = heat_index(h, i, j, k);
                        qn[d][c] = heat_flux_new(h, f, d, c);
                    }
            for (int s = 0; s < 2; ++s) {
                if (h->bc_kind[d][s] != OR_HEAT_FLUX) continue;
                const int face = s == 0 ? 1 : h->n[d] + 1;
                for (int k = kz0; k < kz0 + hz; ++k)
                    for (int j = 1; j <= hy; ++j)
                        for (int i = 1; i <= hx; ++i) {
                            const int idx = d == 0 ? i : (d == 1 ? j : k);
                            if (idx == face) qn[d][heat_index(h, i, j, k)] = h->bc_val[d][s];
                        }
            }
        }
        for (int d = 0; d < nd; ++d) memcpy(f->q[d], qn[d], sizeof(double) * total);
        /* update_T! (kernels.jl:19-29), then the value BCs */
        for (int k = kz0; k < kz0 + nz; ++k)
            for (int j = 1; j <= h->n[1]; ++j)
                for (int i = 1; i <= h->n[0]; ++i) {
                    const long c = heat_index(h, i, j, k);
                    double divq = (f->q[0][c + 1] - f->q[0][c]) / h->dx[0];
                    divq = divq + (f->q[1][c + h->sy] - f->q[1][c]) / h->dx[1];
                    if (nd == 3) divq = divq + (f->q[2][c + h->sz] - f->q[2][c]) / h->dx[2];
                    const double dTdt = (f->T[c] - f->To[c]) / dt;
                    f->T[c] = f->T[c] - (dTdt + divq) * h->dtT;
                }
        or_heat_bc(h, f);
    }
    for (int d = 0; d < nd; ++d) free(qn[d]);
}

static double nanmax(double a, double b) { return (a != a || b != b) ? NAN : (a > b ? a : b); }

void or_heat_residual_norms(const or_heat* h, const or_heat_fields* f, double dt, double out[2])
{
    const int nd = h->ndims;
    const int kz0 = nd == 3 ? 1 : 0;
    double rT = 0.0, rq = 0.0;
    for (int k = kz0; k < kz0 + (nd == 3 ? h->n[2] : 1); ++k)
        for (int j = 1; j <= h->n[1]; ++j)
            for (int i = 1; i <= h->n[0]; ++i) {
                const long c = heat_index(h, i, j, k);
                double divq = (f->q[0][c + 1] - f->q[0][c]) / h->dx[0];
                divq = divq + (f->q[1][c + h->sy] - f->q[1][c]) / h->dx[1];
                if (nd == 3) divq = divq + (f->q[2][c + h->sz] - f->q[2][c]) / h->dx[2];
                rT = nanmax(rT, fabs((f->T[c] - f->To[c]) / dt + divq));
            }
    for (int d = 0; d < nd; ++d) {
        const int hx = h->n[0] + (d == 0), hy = h->n[1] + (d == 1), hz = nd == 3 ? h->n[2] + (d == 2) : 1;
        for (int k = kz0; k < kz0 + hz; ++k)
            for (int j = 1; j <= hy; ++j)
                for (int i = 1; i <= hx; ++i) {
                    const int idx = d == 0 ? i : (d == 1 ? j : k);
                    if ((idx == 1 && h->bc_kind[d][0] == OR_HEAT_FLUX) ||
                        (idx == h->n[d] + 1 && h->bc_kind[d][1] == OR_HEAT_FLUX))
                        continue;
                    const long c = heat_index(h, i, j, k);
                    rq = nanmax(rq, fabs(f->q[d][c] + h->lam * (f->T[c] - f->T[c - heat_stride(h, d)]) / h->dx[d]));
                }
    }
    out[0] = rT;
    out[1] = rq;
}
