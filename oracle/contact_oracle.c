/*
 * CPU ORACLE -- TEST INFRASTRUCTURE ONLY (see sgipc_oracle.h).
 *
 * Contact-pair generation and the barrier Hessians of every pair type,
 * restated from the reference:
 *   - candidate filters, distance types and pair encoding:
 *     StiffGIPC/mlbvh.cu:229-533 (_d_PP/_d_PE/_d_PT/_d_EE, _dType_PT,
 *     _dType_EE), :535-1060 (_checkPTintersection, _checkEEintersection with
 *     smooth = false), :1306-1407 (_selfQuery_vf), :1509-1632
 *     (_selfQuery_ee), :33-42 (overlap with gap sqrt(dHat));
 *   - barrier gradient and Hessian, RANK 2 / NEWF branches of
 *     _calBarrierGradientAndHessian: StiffGIPC/GIPC.cu:3176-3417 (EE),
 *     :3932-4175 (PP), :4436-4731 (PE), :4733-4984 (PT); the reference's
 *     pFpx_*2 (GIPC_PDerivative.cuh:1861-2394) only enter through the column
 *     that holds the gradient of s = dist / sqrt(dHat), written here as the
 *     gradient of the point-point, point-line, point-plane and line-line
 *     distances (closest-point / envelope form);
 *   - arena layout: [PT + EE pairs x 16 | PE x 9 | PP x 4] blocks at
 *     matIndex * 16, n4 * 16 + matIndex * 9, n4 * 16 + n3 * 9 + matIndex * 4
 *     (GIPC.cu:3414,4172,4728,4981), lower blocks moved to (col, row) and
 *     zeroed (GIPC.cu:10344-10362).
 * The reference finds pairs with an LBVH and appends them with atomics (an
 * order that changes run to run); here every candidate is tested by brute
 * force and pairs come out in a fixed order: the point-face query's pairs by
 * (surface-vertex slot, face), then the edge-edge query's by (edge, edge).
 * MatIndex counts each class in that order.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "sgipc_oracle.h"

static void sub3(const double* a, const double* b, double* o)
{
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}
static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void cross3(const double* a, const double* b, double* o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static double det3c(const double* c0, const double* c1, const double* c2)
{
    /* determinant of the matrix with columns c0, c1, c2 */
    double t[3];
    cross3(c1, c2, t);
    return dot3(c0, t);
}

/* squared distances (mlbvh.cu:229-271) */
static double dd_pp(const double* a, const double* b)
{
    double d[3];
    sub3(a, b, d);
    return dot3(d, d);
}
static double dd_pe(const double* p, const double* a, const double* b)
{
    double u[3], v[3], c[3], e[3];
    sub3(a, p, u);
    sub3(b, p, v);
    cross3(u, v, c);
    sub3(b, a, e);
    return dot3(c, c) / dot3(e, e);
}
static double dd_pt(const double* p, const double* a, const double* b, const double* c)
{
    double e1[3], e2[3], n[3], w[3];
    sub3(b, a, e1);
    sub3(c, a, e2);
    cross3(e1, e2, n);
    sub3(p, a, w);
    double s = dot3(w, n);
    return s * s / dot3(n, n);
}
static double dd_ee(const double* a0, const double* a1, const double* b0, const double* b1)
{
    double ea[3], eb[3], n[3], w[3];
    sub3(a1, a0, ea);
    sub3(b1, b0, eb);
    cross3(ea, eb, n);
    sub3(b0, a0, w);
    double s = dot3(w, n);
    return s * s / dot3(n, n);
}

/* which feature of triangle (t0, t1, t2) is closest to p: 0..2 a vertex,
 * 3 edge t0t1, 4 edge t1t2, 5 edge t2t0, 6 the interior (mlbvh.cu:288-447):
 * for each edge, p's coordinates in the frame (edge, in-plane normal of the
 * edge, triangle normal) */
static void edge_params(const double* p, const double* a, const double* b, const double* nv, double* s, double* t)
{
    double e[3], m[3], w[3];
    sub3(b, a, e);
    cross3(e, nv, m);
    sub3(p, a, w);
    double D = det3c(e, m, nv);
    *s = det3c(w, m, nv) / D;
    *t = det3c(e, w, nv) / D;
}
static int dtype_pt(const double* p, const double* t0, const double* t1, const double* t2)
{
    double e1[3], e2[3], nv[3];
    sub3(t1, t0, e1);
    sub3(t2, t0, e2);
    cross3(e1, e2, nv);
    double s0, u0, s1, u1, s2, u2;
    edge_params(p, t0, t1, nv, &s0, &u0);
    if (s0 > 0 && s0 < 1 && u0 >= 0) return 3;
    edge_params(p, t1, t2, nv, &s1, &u1);
    if (s1 > 0.0 && s1 < 1.0 && u1 >= 0.0) return 4;
    edge_params(p, t2, t0, nv, &s2, &u2);
    if (s2 > 0.0 && s2 < 1.0 && u2 >= 0.0) return 5;
    if (s0 <= 0.0 && s2 >= 1.0) return 0;
    if (s1 <= 0.0 && s0 >= 1.0) return 1;
    if (s2 <= 0.0 && s1 >= 1.0) return 2;
    return 6;
}

/* closest features of segments a0a1, b0b1 (mlbvh.cu:449-533): 0 a0-b0,
 * 1 a0-b1, 2 a0-edge b, 3 a1-b0, 4 a1-b1, 5 a1-edge b, 6 b0-edge a,
 * 7 b1-edge a, 8 edge-edge */
static int dtype_ee(const double* a0, const double* a1, const double* b0, const double* b1)
{
    double u[3], v[3], w[3], uv[3];
    sub3(a1, a0, u);
    sub3(b1, b0, v);
    sub3(a0, b0, w);
    double a = dot3(u, u), b = dot3(u, v), c = dot3(v, v), d = dot3(u, w), e = dot3(v, w);
    double D = a * c - b * b, tD = D, sN, tN;
    int res = 8;
    sN = b * e - c * d;
    if (sN <= 0.0) {
        tN = e;
        tD = c;
        res = 2;
    } else if (sN >= D) {
        tN = e + b;
        tD = c;
        res = 5;
    } else {
        tN = a * e - b * d;
        cross3(u, v, uv);
        if (tN > 0.0 && tN < tD && (dot3(w, uv) == 0.0 || dot3(uv, uv) < 1.0e-20 * a * c)) {
            if (sN < D / 2) {
                tN = e;
                tD = c;
                res = 2;
            } else {
                tN = e + b;
                tD = c;
                res = 5;
            }
        }
    }
    if (tN <= 0.0) {
        if (-d <= 0.0) return 0;
        if (-d >= a) return 3;
        return 6;
    }
    if (tN >= tD) {
        if ((-d + b) <= 0.0) return 1;
        if ((-d + b) >= a) return 4;
        return 7;
    }
    return res;
}

/* mollifier threshold of two rest edges (mlbvh.cu:273-286) */
static double eps_x(const double* a0, const double* a1, const double* b0, const double* b1)
{
    return 1e-3 * dd_pp(a0, a1) * dd_pp(b0, b1);
}

typedef struct {
    double lo[3], hi[3];
} Box;
static void box_of(Box* bx, const double* x, const int* v, int k)
{
    for (int d = 0; d < 3; ++d) {
        bx->lo[d] = 1e32;
        bx->hi[d] = -1e32;
    }
    for (int i = 0; i < k; ++i)
        for (int d = 0; d < 3; ++d) {
            double c = x[3 * (size_t)v[i] + d];
            if (c < bx->lo[d]) bx->lo[d] = c;
            if (c > bx->hi[d]) bx->hi[d] = c;
        }
}
/* mlbvh.cu:33-42: separated by at least gap on some axis -> no overlap */
static int overlap(const Box* a, const Box* b, double gap)
{
    for (int d = 0; d < 3; ++d)
        if ((b->lo[d] - a->hi[d]) >= gap || (a->lo[d] - b->hi[d]) >= gap) return 0;
    return 1;
}

static void emit(int* pairs, int* cls, int cap, int* n, int x, int y, int z, int w, int c)
{
    if (*n < cap) {
        int* o = pairs + 4 * (size_t)(*n);
        o[0] = x;
        o[1] = y;
        o[2] = z;
        o[3] = w;
        cls[*n] = c;
    }
    ++*n;
}

/* _checkPTintersection (mlbvh.cu:535-651): encodings (-p-1, q, -1, -1) PP,
 * (-p-1, a, b, -1) PE, (-p-1, t0, t1, t2) PT */
static void check_pt(const double* x, int p, int t0, int t1, int t2, double dhat_sq, int* pairs, int* cls, int cap,
                     int* n)
{
    const double *P = x + 3 * (size_t)p, *A = x + 3 * (size_t)t0, *B = x + 3 * (size_t)t1, *C = x + 3 * (size_t)t2;
    int t = dtype_pt(P, A, B, C);
    double d = 100.0;
    switch (t) {
    case 0: d = dd_pp(P, A); if (d < dhat_sq) emit(pairs, cls, cap, n, -p - 1, t0, -1, -1, 2); break;
    case 1: d = dd_pp(P, B); if (d < dhat_sq) emit(pairs, cls, cap, n, -p - 1, t1, -1, -1, 2); break;
    case 2: d = dd_pp(P, C); if (d < dhat_sq) emit(pairs, cls, cap, n, -p - 1, t2, -1, -1, 2); break;
    case 3: d = dd_pe(P, A, B); if (d < dhat_sq) emit(pairs, cls, cap, n, -p - 1, t0, t1, -1, 3); break;
    case 4: d = dd_pe(P, B, C); if (d < dhat_sq) emit(pairs, cls, cap, n, -p - 1, t1, t2, -1, 3); break;
    case 5: d = dd_pe(P, C, A); if (d < dhat_sq) emit(pairs, cls, cap, n, -p - 1, t2, t0, -1, 3); break;
    default: d = dd_pt(P, A, B, C); if (d < dhat_sq) emit(pairs, cls, cap, n, -p - 1, t0, t1, t2, 4); break;
    }
}

/* _checkEEintersection with smooth = false (mlbvh.cu:685-1060): point-point
 * and point-edge closest features are encoded like the point-face ones with
 * w = -obj - 2 when the edges are nearly parallel at rest, else -1; the
 * edge-edge case keeps (a0, a1, b0, b1) */
static void check_ee(const double* x, const double* rest, int a0, int a1, int b0, int b1, int obj, double dhat_sq,
                     int* pairs, int* cls, int cap, int* n)
{
    const double *A0 = x + 3 * (size_t)a0, *A1 = x + 3 * (size_t)a1, *B0 = x + 3 * (size_t)b0,
                 *B1 = x + 3 * (size_t)b1;
    int t = dtype_ee(A0, A1, B0, B1);
    double ua[3], ub[3], c[3];
    double d;
    int add_e;
    /* cross of (a0 - a1) x (b0 - b1) against the rest threshold; cases 6, 7
     * take the edges in the other order (same value) */
    sub3(A0, A1, ua);
    sub3(B0, B1, ub);
    cross3(ua, ub, c);
    double cr = dot3(c, c);
    double ex = (t == 6 || t == 7)
                    ? eps_x(rest + 3 * (size_t)b0, rest + 3 * (size_t)b1, rest + 3 * (size_t)a0, rest + 3 * (size_t)a1)
                    : eps_x(rest + 3 * (size_t)a0, rest + 3 * (size_t)a1, rest + 3 * (size_t)b0, rest + 3 * (size_t)b1);
    add_e = (cr < ex) ? -obj - 2 : -1;
    switch (t) {
    case 0: d = dd_pp(A0, B0); if (d < dhat_sq) emit(pairs, cls, cap, n, -a0 - 1, b0, -1, add_e, 2); break;
    case 1: d = dd_pp(A0, B1); if (d < dhat_sq) emit(pairs, cls, cap, n, -a0 - 1, b1, -1, add_e, 2); break;
    case 2: d = dd_pe(A0, B0, B1); if (d < dhat_sq) emit(pairs, cls, cap, n, -a0 - 1, b0, b1, add_e, 3); break;
    case 3: d = dd_pp(A1, B0); if (d < dhat_sq) emit(pairs, cls, cap, n, -a1 - 1, b0, -1, add_e, 2); break;
    case 4: d = dd_pp(A1, B1); if (d < dhat_sq) emit(pairs, cls, cap, n, -a1 - 1, b1, -1, add_e, 2); break;
    case 5: d = dd_pe(A1, B0, B1); if (d < dhat_sq) emit(pairs, cls, cap, n, -a1 - 1, b0, b1, add_e, 3); break;
    case 6: d = dd_pe(B0, A0, A1); if (d < dhat_sq) emit(pairs, cls, cap, n, -b0 - 1, a0, a1, add_e, 3); break;
    case 7: d = dd_pe(B1, A0, A1); if (d < dhat_sq) emit(pairs, cls, cap, n, -b1 - 1, a0, a1, add_e, 3); break;
    default: d = dd_ee(A0, A1, B0, B1); if (d < dhat_sq) emit(pairs, cls, cap, n, a0, a1, b0, b1, 4); break;
    }
}

static int skip_all_fixed(const int* btype, const int* v, int k)
{
    if (!btype) return 0;
    for (int i = 0; i < k; ++i)
        if (btype[v[i]] < 2) return 0;
    return 1;
}

int orc_contact_detect(int n_verts, const double* x, const double* rest, int n_faces, const int* faces, int n_edges,
                       const int* edges, int n_surf, const int* surf, const int* body_id, const int* btype,
                       double dhat_sq, int capacity, int* pairs, int* mat_index, int* counts)
{
    (void)n_verts;
    const double gap = sqrt(dhat_sq);
    int* cls = (int*)malloc(sizeof(int) * (size_t)(capacity > 0 ? capacity : 1));
    int n = 0;
    Box* fb = (Box*)malloc(sizeof(Box) * (size_t)(n_faces > 0 ? n_faces : 1));
    for (int f = 0; f < n_faces; ++f) box_of(&fb[f], x, faces + 3 * (size_t)f, 3);
    /* point-face query (mlbvh.cu:1306-1407) */
    for (int q = 0; q < n_surf; ++q) {
        int p = surf[q];
        Box pb;
        box_of(&pb, x, &p, 1);
        for (int f = 0; f < n_faces; ++f) {
            const int* tv = faces + 3 * (size_t)f;
            if (!overlap(&pb, &fb[f], gap)) continue;
            if (body_id && body_id[p] == body_id[tv[0]] && body_id[p] != -1) continue;
            if (p == tv[0] || p == tv[1] || p == tv[2]) continue;
            int all[4] = {p, tv[0], tv[1], tv[2]};
            if (skip_all_fixed(btype, all, 4)) continue;
            check_pt(x, p, tv[0], tv[1], tv[2], dhat_sq, pairs, cls, capacity, &n);
        }
    }
    free(fb);
    /* edge-edge query (mlbvh.cu:1509-1632): each pair once, the larger edge id
     * as the found object */
    Box* eb = (Box*)malloc(sizeof(Box) * (size_t)(n_edges > 0 ? n_edges : 1));
    for (int e = 0; e < n_edges; ++e) box_of(&eb[e], x, edges + 2 * (size_t)e, 2);
    for (int e = 0; e < n_edges; ++e) {
        const int* ev = edges + 2 * (size_t)e;
        for (int o = e + 1; o < n_edges; ++o) {
            const int* ov = edges + 2 * (size_t)o;
            if (!overlap(&eb[e], &eb[o], gap)) continue;
            if (body_id && body_id[ev[0]] == body_id[ov[0]] && body_id[ev[0]] != -1) continue;
            if (ev[0] == ov[0] || ev[0] == ov[1] || ev[1] == ov[0] || ev[1] == ov[1]) continue;
            int all[4] = {ev[0], ev[1], ov[0], ov[1]};
            if (skip_all_fixed(btype, all, 4)) continue;
            check_ee(x, rest, ev[0], ev[1], ov[0], ov[1], o, dhat_sq, pairs, cls, capacity, &n);
        }
    }
    free(eb);
    counts[0] = n;
    counts[1] = 0;
    counts[2] = counts[3] = counts[4] = 0;
    /* class counts and MatIndex of the pairs written (all of them unless the
     * capacity was short) */
    int m = n < capacity ? n : capacity;
    for (int i = 0; i < m; ++i) mat_index[i] = counts[cls[i]]++;
    free(cls);
    return n > capacity ? -1 : 0;
}

/* gradient of s = dist / sqrt(dHat) over the pair's k vertices (k = 2, 3, 4)
 * and dist^2; returns k.  Types: 2 PP, 3 PE, 4 PT, 5 EE. */
static int pair_geometry(const double* x, const int* pr, int* vid, double* g, double* d2, int* type)
{
    double P[4][3];
    int k;
    if (pr[0] >= 0) {
        if (pr[3] < 0) return -1; /* smooth-encoded pair: not produced with smooth = false */
        vid[0] = pr[0]; vid[1] = pr[1]; vid[2] = pr[2]; vid[3] = pr[3];
        k = 4;
        *type = 5;
    } else {
        vid[0] = -pr[0] - 1;
        if (pr[1] < 0) return -1;
        if (pr[2] < 0) {
            vid[1] = pr[1];
            k = 2;
            *type = 2;
        } else if (pr[3] < 0) {
            vid[1] = pr[1];
            vid[2] = pr[2];
            k = 3;
            *type = 3;
        } else {
            vid[1] = pr[1]; vid[2] = pr[2]; vid[3] = pr[3];
            k = 4;
            *type = 4;
        }
    }
    for (int a = 0; a < k; ++a)
        for (int d = 0; d < 3; ++d) P[a][d] = x[3 * (size_t)vid[a] + d];
    double r[3], coef[4];
    if (*type == 2) {
        sub3(P[0], P[1], r);
        coef[0] = 1.0;
        coef[1] = -1.0;
    } else if (*type == 3) {
        /* foot of p on the line a + t (b - a) */
        double e[3], u[3];
        sub3(P[2], P[1], e);
        sub3(P[0], P[1], u);
        double t = dot3(u, e) / dot3(e, e);
        for (int d = 0; d < 3; ++d) r[d] = u[d] - t * e[d];
        coef[0] = 1.0;
        coef[1] = orc_get_fault() == 5 ? -t : -(1.0 - t); /* fault 5: planted wrong weight (tests) */
        coef[2] = -t;
    } else if (*type == 4) {
        /* foot of p on the plane: barycentric (b0, b1, b2) */
        double e1[3], e2[3], w[3], nv[3], c1[3], c2[3];
        sub3(P[2], P[1], e1);
        sub3(P[3], P[1], e2);
        sub3(P[0], P[1], w);
        cross3(e1, e2, nv);
        double nn = dot3(nv, nv);
        cross3(w, e2, c1);
        cross3(e1, w, c2);
        double b1 = dot3(c1, nv) / nn, b2 = dot3(c2, nv) / nn, b0 = 1.0 - b1 - b2;
        double s = dot3(w, nv) / nn;
        for (int d = 0; d < 3; ++d) r[d] = s * nv[d];
        coef[0] = 1.0;
        coef[1] = -b0;
        coef[2] = -b1;
        coef[3] = -b2;
    } else {
        /* closest points of the lines a0 + s ea and b0 + t eb */
        double ea[3], eb[3], w0[3];
        sub3(P[1], P[0], ea);
        sub3(P[3], P[2], eb);
        sub3(P[0], P[2], w0);
        double a = dot3(ea, ea), b = dot3(ea, eb), c = dot3(eb, eb), dd = dot3(ea, w0), e = dot3(eb, w0);
        double D = a * c - b * b;
        double s = (b * e - c * dd) / D, t = (a * e - b * dd) / D;
        for (int d = 0; d < 3; ++d) r[d] = w0[d] + s * ea[d] - t * eb[d];
        coef[0] = 1.0 - s;
        coef[1] = s;
        coef[2] = -(1.0 - t);
        coef[3] = -t;
    }
    double rr = dot3(r, r);
    double dis = sqrt(rr);
    /* the reference's distance values (mlbvh.cu:229-258) */
    if (*type == 2) *d2 = dd_pp(P[0], P[1]);
    else if (*type == 3) *d2 = dd_pe(P[0], P[1], P[2]);
    else if (*type == 4) *d2 = dd_pt(P[0], P[1], P[2], P[3]);
    else *d2 = dd_ee(P[0], P[1], P[2], P[3]);
    for (int a = 0; a < k; ++a)
        for (int d = 0; d < 3; ++d) g[3 * a + d] = coef[a] * r[d] / dis;
    return k;
}

int orc_barrier(int n_pairs, const double* x, const int* pairs, const int* mat_index, int n4, int n3, double dhat_sq,
                double kappa, int hess_offset, double* grad, double* tvals, int* trows, int* tcols)
{
    const double gass = 1e-6;
    const double sq = sqrt(dhat_sq);
    int bad = 0;
    for (int p = 0; p < n_pairs; ++p) {
        int vid[4], type;
        double g[12], d2;
        int k = pair_geometry(x, pairs + 4 * (size_t)p, vid, g, &d2, &type);
        if (k < 0) {
            bad = 1;
            continue;
        }
        for (int i = 0; i < 3 * k; ++i) g[i] /= sq;
        double I5 = d2 / dhat_sq;
        double lam0 = orc_barrier_lambda0(d2 < gass * dhat_sq ? gass : I5, dhat_sq, kappa);
        if (grad) {
            double L = log(I5);
            double s = sqrt(d2) / sq;
            double pk = 2.0 * (2.0 * kappa * dhat_sq * dhat_sq * L * (I5 - 1.0) * (I5 + I5 * L - 1.0)) / I5;
            for (int a = 0; a < k; ++a)
                for (int d = 0; d < 3; ++d) grad[3 * (size_t)vid[a] + d] += g[3 * a + d] * s * pk;
        }
        long base;
        int mi = mat_index[p];
        if (k == 4) base = 16L * mi;
        else if (k == 3) base = 16L * n4 + 9L * mi;
        else base = 16L * n4 + 9L * n3 + 4L * mi;
        for (int ii = 0; ii < k; ++ii)
            for (int jj = 0; jj < k; ++jj) {
                long kk = base + ii * k + jj;
                int row = vid[ii] + hess_offset, col = vid[jj] + hess_offset;
                double* v = tvals + 9 * kk;
                if (row <= col) {
                    trows[kk] = row;
                    tcols[kk] = col;
                    for (int r = 0; r < 3; ++r)
                        for (int c = 0; c < 3; ++c) v[c * 3 + r] = lam0 * g[3 * ii + r] * g[3 * jj + c];
                } else {
                    trows[kk] = col;
                    tcols[kk] = row;
                    for (int e = 0; e < 9; ++e) v[e] = 0.0;
                }
            }
    }
    return bad ? -1 : 0;
}

/* barrier energy of the pairs (the FD pin of the gradient) */
double orc_barrier_energy_pairs(int n_pairs, const double* x, const int* pairs, double dhat_sq, double kappa)
{
    double e = 0.0;
    for (int p = 0; p < n_pairs; ++p) {
        int vid[4], type;
        double g[12], d2;
        if (pair_geometry(x, pairs + 4 * (size_t)p, vid, g, &d2, &type) < 0) continue;
        e += orc_barrier_energy(d2 / dhat_sq, dhat_sq, kappa);
    }
    return e;
}

/* ------------------------------------------------------------------------- */
/* Feasible step: additive CCD of the candidate pairs (StiffGIPC/ACCD.cu:      */
/* 317-534, eta = 1 - slackness, thickness 0, the direction -moveDir as in    */
/* _reduct_min_selfTimeStep_to_double, GIPC.cu:7095-7198), candidates by the  */
/* swept-box full-CCD queries (mlbvh.cu:1408-1508,1634-1774) with the point-  */
/* face filter of _checkPTintersection_fullCCD (:653-683); the ground step     */
/* (GIPC.cu:6945-7020) and the tet injectivity step (GIPC.cu:1130-1300,       */
/* 7022-7094).                                                                */
/* ------------------------------------------------------------------------- */
static double pt_dist_unclassified(const double* p, const double* a, const double* b, const double* c)
{
    switch (dtype_pt(p, a, b, c)) {
    case 0: return dd_pp(p, a);
    case 1: return dd_pp(p, b);
    case 2: return dd_pp(p, c);
    case 3: return dd_pe(p, a, b);
    case 4: return dd_pe(p, b, c);
    case 5: return dd_pe(p, c, a);
    default: return dd_pt(p, a, b, c);
    }
}
static double ee_dist_unclassified(const double* a0, const double* a1, const double* b0, const double* b1)
{
    switch (dtype_ee(a0, a1, b0, b1)) {
    case 0: return dd_pp(a0, b0);
    case 1: return dd_pp(a0, b1);
    case 2: return dd_pe(a0, b0, b1);
    case 3: return dd_pp(a1, b0);
    case 4: return dd_pp(a1, b1);
    case 5: return dd_pe(a1, b0, b1);
    case 6: return dd_pe(b0, a0, a1);
    case 7: return dd_pe(b1, a0, a1);
    default: return dd_ee(a0, a1, b0, b1);
    }
}

/* conservative advancement of 4 points moving with d[4][3] (vertex 0..3), the
 * distance function f (point-triangle or edge-edge); returns the time of
 * contact in [0, 1] */
static double accd(double P[4][3], double D[4][3], int ee, double eta, double th)
{
    double mov[3];
    for (int d = 0; d < 3; ++d) mov[d] = -0.25 * (D[0][d] + D[1][d] + D[2][d] + D[3][d]);
    for (int a = 0; a < 4; ++a)
        for (int d = 0; d < 3; ++d) D[a][d] += mov[d];
    double maxd;
    if (ee)
        maxd = sqrt(fmax(dot3(D[0], D[0]), dot3(D[1], D[1]))) + sqrt(fmax(dot3(D[2], D[2]), dot3(D[3], D[3])));
    else
        maxd = sqrt(dot3(D[0], D[0])) + sqrt(fmax(dot3(D[1], D[1]), fmax(dot3(D[2], D[2]), dot3(D[3], D[3]))));
    if (maxd == 0.0) return 1.0;
#define ORC_DIST2(out)                                                                    \
    do {                                                                                  \
        if (ee) {                                                                         \
            out = ee_dist_unclassified(P[0], P[1], P[2], P[3]);                           \
            if (out - th * th <= 0.0)                                                     \
                out = fmin(fmin(dd_pp(P[0], P[2]), dd_pp(P[0], P[3])),                    \
                           fmin(dd_pp(P[1], P[2]), dd_pp(P[1], P[3])));                   \
        } else {                                                                          \
            out = pt_dist_unclassified(P[0], P[1], P[2], P[3]);                           \
        }                                                                                 \
    } while (0)
    double d2;
    ORC_DIST2(d2);
    double dc = sqrt(d2);
    double gap = eta * (d2 - th * th) / (dc + th);
    double toc = 0.0;
    int count = 0;
    for (;;) {
        if (++count > 50000) return toc;
        double tlb = (1.0 - eta) * (d2 - th * th) / ((dc + th) * maxd);
        for (int a = 0; a < 4; ++a)
            for (int d = 0; d < 3; ++d) P[a][d] += D[a][d] * tlb;
        ORC_DIST2(d2);
        dc = sqrt(d2);
        if (toc != 0.0 && (d2 - th * th) / (dc + th) < gap) break;
        toc += tlb;
        if (toc > 1.0) return 1.0;
    }
#undef ORC_DIST2
    return toc;
}

static void swept_box(Box* bx, const double* x, const double* dir, double alpha, const int* v, int k)
{
    box_of(bx, x, v, k);
    for (int i = 0; i < k; ++i)
        for (int d = 0; d < 3; ++d) {
            double c = x[3 * (size_t)v[i] + d] - dir[3 * (size_t)v[i] + d] * alpha;
            if (c < bx->lo[d]) bx->lo[d] = c;
            if (c > bx->hi[d]) bx->hi[d] = c;
        }
}

double orc_ccd_self_step(int n_verts, const double* x, const double* dir, double alpha, int n_faces, const int* faces,
                         int n_edges, const int* edges, int n_surf, const int* surf, const int* body_id,
                         const int* btype, double dhat_sq, double slackness, int* n_candidates)
{
    (void)n_verts;
    const double gap = sqrt(dhat_sq), eta = 1.0 - slackness;
    double tmax = 1.0; /* max of 1 / toc */
    int cand = 0;
    Box* fb = (Box*)malloc(sizeof(Box) * (size_t)(n_faces > 0 ? n_faces : 1));
    for (int f = 0; f < n_faces; ++f) swept_box(&fb[f], x, dir, alpha, faces + 3 * (size_t)f, 3);
    for (int q = 0; q < n_surf; ++q) {
        int p = surf[q];
        Box pb;
        swept_box(&pb, x, dir, alpha, &p, 1);
        for (int f = 0; f < n_faces; ++f) {
            const int* tv = faces + 3 * (size_t)f;
            if (!overlap(&pb, &fb[f], gap)) continue;
            if (body_id && body_id[p] == body_id[tv[0]] && body_id[p] != -1) continue;
            if (p == tv[0] || p == tv[1] || p == tv[2]) continue;
            int all[4] = {p, tv[0], tv[1], tv[2]};
            if (skip_all_fixed(btype, all, 4)) continue;
            const double *P0 = x + 3 * (size_t)p, *A = x + 3 * (size_t)tv[0], *B = x + 3 * (size_t)tv[1],
                         *C = x + 3 * (size_t)tv[2];
            double e1[3], e2[3], nv[3], w[3];
            sub3(B, A, e1);
            sub3(C, A, e2);
            cross3(e1, e2, nv);
            sub3(P0, A, w);
            if (dtype_pt(P0, A, B, C) == 6 && dot3(nv, w) < 0) continue;
            ++cand;
            double P[4][3], D[4][3];
            for (int a = 0; a < 4; ++a)
                for (int d = 0; d < 3; ++d) {
                    P[a][d] = x[3 * (size_t)all[a] + d];
                    D[a][d] = -dir[3 * (size_t)all[a] + d];
                }
            double t = 1.0 / accd(P, D, 0, eta, 0.0);
            if (t > tmax) tmax = t;
        }
    }
    free(fb);
    Box* eb = (Box*)malloc(sizeof(Box) * (size_t)(n_edges > 0 ? n_edges : 1));
    for (int e = 0; e < n_edges; ++e) swept_box(&eb[e], x, dir, alpha, edges + 2 * (size_t)e, 2);
    for (int e = 0; e < n_edges; ++e) {
        const int* ev = edges + 2 * (size_t)e;
        for (int o = e + 1; o < n_edges; ++o) {
            const int* ov = edges + 2 * (size_t)o;
            if (!overlap(&eb[e], &eb[o], gap)) continue;
            if (body_id && body_id[ev[0]] == body_id[ov[0]] && body_id[ev[0]] != -1) continue;
            int all[4] = {ev[0], ev[1], ov[0], ov[1]};
            if (skip_all_fixed(btype, all, 4)) continue;
            if (ev[0] == ov[0] || ev[0] == ov[1] || ev[1] == ov[0] || ev[1] == ov[1]) continue;
            ++cand;
            double P[4][3], D[4][3];
            for (int a = 0; a < 4; ++a)
                for (int d = 0; d < 3; ++d) {
                    P[a][d] = x[3 * (size_t)all[a] + d];
                    D[a][d] = -dir[3 * (size_t)all[a] + d];
                }
            double t = 1.0 / accd(P, D, 1, eta, 0.0);
            if (t > tmax) tmax = t;
        }
    }
    free(eb);
    if (n_candidates) *n_candidates = cand;
    return 1.0 / tmax;
}

double orc_ground_step(int n_surf, const double* x, const int* surf, const double* normal, double offset,
                       const double* dir, double slackness)
{
    double tmax = 1.0;
    for (int q = 0; q < n_surf; ++q) {
        const double* xv = x + 3 * (size_t)surf[q];
        const double* dv = dir + 3 * (size_t)surf[q];
        double coef = dot3(normal, dv);
        if (coef > 0.0) {
            double dist = dot3(normal, xv) - offset;
            double t = coef / (dist * slackness);
            if (t > tmax) tmax = t;
        }
    }
    return 1.0 / tmax;
}

/* smallest t > 0 with det(Ds(x - t dir)) = slackness det(Ds(x)): the cubic
 * c3 t^3 + c2 t^2 + c1 t + (1 - slackness) c0 with c0 = det D0, c3 = det D1
 * (D1 the edge matrix of -dir) and c1, c2 the mixed determinants; below
 * error_rate in |c3| (and |c2|) the reference's quadratic / linear branches */
static double det_cols3(const double* a, const double* b, const double* c)
{
    double t[3];
    cross3(b, c, t);
    return dot3(a, t);
}
double orc_injective_step(int n_tets, const double* x, const int* tets, const double* dir, double slackness,
                          double error_rate)
{
    double tmax = 1.0;
    for (int e = 0; e < n_tets; ++e) {
        const int* t = tets + 4 * (size_t)e;
        double A[3][3], Bm[3][3];
        for (int c = 0; c < 3; ++c)
            for (int d = 0; d < 3; ++d) {
                A[c][d] = x[3 * (size_t)t[c + 1] + d] - x[3 * (size_t)t[0] + d];
                Bm[c][d] = -dir[3 * (size_t)t[c + 1] + d] + dir[3 * (size_t)t[0] + d];
            }
        double c0 = det_cols3(A[0], A[1], A[2]), c3 = det_cols3(Bm[0], Bm[1], Bm[2]);
        double c1 = det_cols3(Bm[0], A[1], A[2]) + det_cols3(A[0], Bm[1], A[2]) + det_cols3(A[0], A[1], Bm[2]);
        double c2 = det_cols3(A[0], Bm[1], Bm[2]) + det_cols3(Bm[0], A[1], Bm[2]) + det_cols3(Bm[0], Bm[1], A[2]);
        double k0 = (1.0 - slackness) * c0;
        double ts = 1.0;
        if (fabs(c3) <= error_rate) {
            if (fabs(c2) <= error_rate) {
                ts = -k0 / c1;
            } else {
                double disc = c1 * c1 - 4.0 * c2 * k0;
                if (disc > 0) {
                    ts = (-c1 - sqrt(disc)) / (2.0 * c2);
                    if (ts < 0) ts = (-c1 + sqrt(disc)) / (2.0 * c2);
                } else {
                    ts = 1.0;
                }
            }
        } else {
            /* smallest positive real root of the cubic: bracket by sampling,
             * then bisect (the reference leaves this case undefined) */
            double best = -1.0;
            const int S = 256;
            double tlo = 0.0, glo = k0;
            for (int s = 1; s <= S && best < 0; ++s) {
                double th = 4.0 * s / S;
                double gh = ((c3 * th + c2) * th + c1) * th + k0;
                if ((glo > 0) != (gh > 0) || gh == 0.0) {
                    double lo = tlo, hi = th;
                    for (int it = 0; it < 100; ++it) {
                        double m = 0.5 * (lo + hi), gm = ((c3 * m + c2) * m + c1) * m + k0;
                        if ((gm > 0) == (glo > 0)) lo = m; else hi = m;
                    }
                    best = 0.5 * (lo + hi);
                }
                tlo = th;
                glo = gh;
            }
            ts = best > 0 ? best : 1.0;
        }
        if (ts <= 0) ts = 1.0;
        double inv = 1.0 / ts;
        if (inv > tmax) tmax = inv;
    }
    return 1.0 / tmax;
}
