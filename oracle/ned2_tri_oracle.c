/*
 * CPU ORACLE -- test infrastructure only.
 *
 * Plain-C restatement of the reference's second-order Nedelec triangle
 * kernels used for Robin (port / absorbing) boundary terms.  Checker for the
 * HIP boundary kernel of emerge_amd; never loaded by the product path.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>

typedef double complex cplx;

static const double FACT[10] = {1, 1, 2, 6, 24, 120, 720, 5040, 40320, 362880};

/* fem/mth/optimized.py:262-271  area_coeff */
static double area_coeff(int a, int b, int c, int d) {
    int k[7] = {0, 0, 0, 0, 0, 0, 0};
    k[a]++; k[b]++; k[c]++; k[d]++;
    int s = k[1] + k[2] + k[3] + k[4] + k[5] + k[6];
    return 2.0 * FACT[k[1]] * FACT[k[2]] * FACT[k[3]] * FACT[k[4]] * FACT[k[5]] * FACT[k[6]] / FACT[s + 2];
}

static double dot2(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1]; }   /* tri.py:32-34 */

/* Shared body of tri.py:575-695 (ned2_tri_stiff_force) and tri.py:697-796
 * (ned2_tri_stiff): the 8x8 Robin block on local 2D vertices (xs, ys) with
 * edge lengths Le[3] (force variant: distances between local vertices).     */
static void tri_block(const double *xs, const double *ys, const double *Le, double Area, cplx coeff, cplx *Bmat) {
    const int lem[2][3] = {{0, 1, 0}, {1, 2, 2}};
    double x1 = xs[0], x2 = xs[1], x3 = xs[2], y1 = ys[0], y2 = ys[1], y3 = ys[2];
    double b1 = y2 - y3, b2 = y3 - y1, b3 = y1 - y2, c1 = x3 - x2, c2 = x1 - x3, c3 = x2 - x1;
    double GL[3][2] = {{b1, c1}, {b2, c2}, {b3, c3}};
    double Ds[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Ds[i][j] = sqrt((xs[i] - xs[j]) * (xs[i] - xs[j]) + (ys[i] - ys[j]) * (ys[i] - ys[j]));
    const int letters[6] = {1, 2, 3, 4, 5, 6};
    int tA = letters[0], tB = letters[1], tC = letters[2];
    const double *GtA = GL[0], *GtB = GL[1], *GtC = GL[2];
    double Lt1 = Ds[2][0], Lt2 = Ds[1][0];
#define AC(a, b, c, d) (area_coeff(a, b, c, d) * Area)
#define BM(r, c) Bmat[(r) * 8 + (c)]
    for (int i = 0; i < 64; i++) Bmat[i] = 0;
    for (int ei = 0; ei < 3; ei++) {
        int ei1 = lem[0][ei], ei2 = lem[1][ei];
        double Li = Le[ei];
        int A = letters[ei1], B = letters[ei2];
        const double *GA = GL[ei1], *GB = GL[ei2];
        for (int ej = 0; ej < 3; ej++) {
            int ej1 = lem[0][ej], ej2 = lem[1][ej];
            double Lj = Le[ej];
            int C = letters[ej1], D = letters[ej2];
            const double *GC = GL[ej1], *GD = GL[ej2];
            BM(ei, ej) += Li * Lj * (AC(A, B, C, D) * dot2(GA, GC) - AC(A, B, C, C) * dot2(GA, GD) - AC(A, A, C, D) * dot2(GB, GC) + AC(A, A, C, C) * dot2(GB, GD));
            BM(ei, ej + 4) += Li * Lj * (AC(A, B, D, D) * dot2(GA, GC) - AC(A, B, C, D) * dot2(GA, GD) - AC(A, A, D, D) * dot2(GB, GC) + AC(A, A, C, D) * dot2(GB, GD));
            BM(ei + 4, ej) += Li * Lj * (AC(B, B, C, D) * dot2(GA, GC) - AC(B, B, C, C) * dot2(GA, GD) - AC(A, B, C, D) * dot2(GB, GC) + AC(A, B, C, C) * dot2(GB, GD));
            BM(ei + 4, ej + 4) += Li * Lj * (AC(B, B, D, D) * dot2(GA, GC) - AC(B, B, C, D) * dot2(GA, GD) - AC(A, B, D, D) * dot2(GB, GC) + AC(A, B, C, D) * dot2(GB, GD));
        }
        BM(ei, 3) += Li * Lt1 * (AC(A, B, tA, tB) * dot2(GA, GtC) - AC(A, B, tB, tC) * dot2(GA, GtA) - AC(A, A, tA, tB) * dot2(GB, GtC) + AC(A, A, tB, tC) * dot2(GB, GtA));
        BM(ei, 7) += Li * Lt2 * (AC(A, B, tB, tC) * dot2(GA, GtA) - AC(A, B, tC, tA) * dot2(GA, GtB) - AC(A, A, tB, tC) * dot2(GB, GtA) + AC(A, A, tC, tA) * dot2(GB, GtB));
        BM(3, ei) += Lt1 * Li * (AC(tA, tB, A, B) * dot2(GA, GtC) - AC(tA, tB, A, A) * dot2(GB, GtC) - AC(tB, tC, A, B) * dot2(GA, GtA) + AC(tB, tC, A, A) * dot2(GB, GtA));
        BM(7, ei) += Lt2 * Li * (AC(tB, tC, A, B) * dot2(GA, GtA) - AC(tB, tC, A, A) * dot2(GB, GtA) - AC(tC, tA, A, B) * dot2(GA, GtB) + AC(tC, tA, A, A) * dot2(GB, GtB));
        BM(ei + 4, 3) += Li * Lt1 * (AC(B, B, tA, tB) * dot2(GA, GtC) - AC(B, B, tB, tC) * dot2(GA, GtA) - AC(A, B, tA, tB) * dot2(GB, GtC) + AC(A, B, tB, tC) * dot2(GB, GtA));
        BM(ei + 4, 7) += Li * Lt2 * (AC(B, B, tB, tC) * dot2(GA, GtA) - AC(B, B, tC, tA) * dot2(GA, GtB) - AC(A, B, tB, tC) * dot2(GB, GtA) + AC(A, B, tC, tA) * dot2(GB, GtB));
        BM(3, ei + 4) += Lt1 * Li * (AC(tA, tB, B, B) * dot2(GA, GtC) - AC(tA, tB, A, B) * dot2(GB, GtC) - AC(tB, tC, B, B) * dot2(GA, GtA) + AC(tB, tC, A, B) * dot2(GB, GtA));
        BM(7, ei + 4) += Lt2 * Li * (AC(tB, tC, B, B) * dot2(GA, GtA) - AC(tB, tC, A, B) * dot2(GB, GtA) - AC(tC, tA, B, B) * dot2(GA, GtB) + AC(tC, tA, A, B) * dot2(GB, GtB));
    }
    BM(3, 3) += Lt1 * Lt1 * (AC(tA, tB, tA, tB) * dot2(GtC, GtC) - AC(tA, tB, tB, tC) * dot2(GtA, GtC) - AC(tB, tC, tA, tB) * dot2(GtA, GtC) + AC(tB, tC, tB, tC) * dot2(GtA, GtA));
    BM(3, 7) += Lt1 * Lt2 * (AC(tA, tB, tB, tC) * dot2(GtA, GtC) - AC(tA, tB, tC, tA) * dot2(GtB, GtC) - AC(tB, tC, tB, tC) * dot2(GtA, GtA) + AC(tB, tC, tC, tA) * dot2(GtA, GtB));
    BM(7, 3) += Lt2 * Lt1 * (AC(tB, tC, tA, tB) * dot2(GtA, GtC) - AC(tB, tC, tB, tC) * dot2(GtA, GtA) - AC(tC, tA, tA, tB) * dot2(GtB, GtC) + AC(tC, tA, tB, tC) * dot2(GtA, GtB));
    BM(7, 7) += Lt2 * Lt2 * (AC(tB, tC, tB, tC) * dot2(GtA, GtA) - AC(tB, tC, tC, tA) * dot2(GtA, GtB) - AC(tC, tA, tB, tC) * dot2(GtA, GtB) + AC(tC, tA, tC, tA) * dot2(GtB, GtB));
    for (int i = 0; i < 64; i++) Bmat[i] *= coeff;
#undef AC
#undef BM
}

/* fem/mth/tri.py:575-695  ned2_tri_stiff_force.
 * lcs (3,3) row-major local vertex coords (only x and y rows used);
 * Uinc (3, nq) row-major local incident field at the nq quadrature points;
 * dpts (4, nq) row-major: weights, then the three barycentric rows.          */
void oracle_ned2_tri_stiff_force(const double *lcs, double gre, double gim, const cplx *Uinc, const double *dpts,
                                 int nq, cplx *Bmat, cplx *bvec) {
    const cplx gamma = gre + I * gim;
    const int lem[2][3] = {{0, 1, 0}, {1, 2, 2}};
    const double *xs = lcs, *ys = lcs + 3;
    double x1 = xs[0], x2 = xs[1], x3 = xs[2], y1 = ys[0], y2 = ys[1], y3 = ys[2];
    double As[3] = {x2 * y3 - y2 * x3, x3 * y1 - y3 * x1, x1 * y2 - y1 * x2};
    double Bs[3] = {y2 - y3, y3 - y1, y1 - y2};
    double Cs[3] = {x3 - x2, x1 - x3, x2 - x1};
    double Ds[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Ds[i][j] = sqrt((xs[i] - xs[j]) * (xs[i] - xs[j]) + (ys[i] - ys[j]) * (ys[i] - ys[j]));
    double Area = 0.5 * fabs((x1 - x3) * (y2 - y1) - (x1 - x2) * (y3 - y1));
    double sv = (x1 - x3) * (y2 - y1) - (x1 - x2) * (y3 - y1);
    double signA = -((sv > 0) - (sv < 0));
    double Le[3] = {Ds[lem[0][0]][lem[1][0]], Ds[lem[0][1]][lem[1][1]], Ds[lem[0][2]][lem[1][2]]};   /* :627 Li = Ds[ei1, ei2] */
    cplx coeff = gamma / ((2 * Area) * (2 * Area));
    tri_block(xs, ys, Le, Area, coeff, Bmat);
    double Lt1 = Ds[2][0], Lt2 = Ds[1][0];
    for (int i = 0; i < 8; i++) bvec[i] = 0;
    const double *Ws = dpts;
    for (int q = 0; q < nq; q++) {
        double x = x1 * dpts[nq + q] + x2 * dpts[2 * nq + q] + x3 * dpts[3 * nq + q];
        double y = y1 * dpts[nq + q] + y2 * dpts[2 * nq + q] + y3 * dpts[3 * nq + q];
        cplx Ux = Uinc[q], Uy = Uinc[nq + q];
        double A2 = 4 * Area * Area, A3 = 8 * Area * Area * Area;
        for (int ei = 0; ei < 3; ei++) {                                  /* :683-690 */
            int ei1 = lem[0][ei], ei2 = lem[1][ei];
            double Li = Le[ei];
            double A1_ = As[ei1], A2_ = As[ei2], B1 = Bs[ei1], B2 = Bs[ei2], C1 = Cs[ei1], C2 = Cs[ei2];
            double l1 = A1_ + B1 * x + C1 * y, l2 = A2_ + B2 * x + C2 * y;
            double Ee1x = Li * (B1 * l2 / A2 - B2 * l1 / A2) * l1 / (2 * Area);
            double Ee1y = Li * (C1 * l2 / A2 - C2 * l1 / A2) * l1 / (2 * Area);
            double Ee2x = Li * (B1 * l2 / A2 - B2 * l1 / A2) * l2 / (2 * Area);
            double Ee2y = Li * (C1 * l2 / A2 - C2 * l1 / A2) * l2 / (2 * Area);
            bvec[ei] += signA * Area * Ws[q] * (Ee1x * Ux + Ee1y * Uy);
            bvec[ei + 4] += signA * Area * Ws[q] * (Ee2x * Ux + Ee2y * Uy);
        }
        double l1 = As[0] + Bs[0] * x + Cs[0] * y, l2 = As[1] + Bs[1] * x + Cs[1] * y, l3 = As[2] + Bs[2] * x + Cs[2] * y;
        double Ef1x = Lt1 * (-Bs[0] * l2 * l3 / A3 + Bs[2] * l1 * l2 / A3);          /* :684-687 */
        double Ef1y = Lt1 * (-Cs[0] * l2 * l3 / A3 + Cs[2] * l1 * l2 / A3);
        double Ef2x = Lt2 * (Bs[0] * l2 * l3 / A3 - Bs[1] * l1 * l3 / A3);
        double Ef2y = Lt2 * (Cs[0] * l2 * l3 / A3 - Cs[1] * l1 * l3 / A3);
        bvec[3] += signA * Area * Ws[q] * (Ef1x * Ux + Ef1y * Uy);
        bvec[7] += signA * Area * Ws[q] * (Ef2x * Ux + Ef2y * Uy);
    }
}

/* fem/mth/tri.py:697-796  ned2_tri_stiff: projects the 3D triangle into its
 * own plane first (:708-721), edge lengths are the mesh's (:740).           */
void oracle_ned2_tri_stiff(const double *verts, const double *edge_lengths, double gre, double gim, cplx *Bmat) {
    const cplx gamma = gre + I * gim;
    const double *xs = verts, *ys = verts + 3, *zs = verts + 6;
    double ax1[3] = {xs[1] - xs[0], ys[1] - ys[0], zs[1] - zs[0]};
    double ax2[3] = {xs[2] - xs[0], ys[2] - ys[0], zs[2] - zs[0]};
    double n1 = sqrt(ax1[0] * ax1[0] + ax1[1] * ax1[1] + ax1[2] * ax1[2]);
    double n2 = sqrt(ax2[0] * ax2[0] + ax2[1] * ax2[1] + ax2[2] * ax2[2]);
    for (int i = 0; i < 3; i++) { ax1[i] /= n1; ax2[i] /= n2; }
    double axn[3] = {ax1[1] * ax2[2] - ax1[2] * ax2[1], ax1[2] * ax2[0] - ax1[0] * ax2[2], ax1[0] * ax2[1] - ax1[1] * ax2[0]};
    double b2[3] = {-(axn[1] * ax1[2] - axn[2] * ax1[1]), -(axn[2] * ax1[0] - axn[0] * ax1[2]), -(axn[0] * ax1[1] - axn[1] * ax1[0])};
    /* basis columns (ax1, b2, axn) are orthogonal, so pinv = transpose scaled
     * by the inverse squared column norms.                                   */
    double c0 = ax1[0] * ax1[0] + ax1[1] * ax1[1] + ax1[2] * ax1[2];
    double c1 = b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2];
    double lx[3], ly[3];
    for (int k = 0; k < 3; k++) {
        lx[k] = (ax1[0] * xs[k] + ax1[1] * ys[k] + ax1[2] * zs[k]) / c0;
        ly[k] = (b2[0] * xs[k] + b2[1] * ys[k] + b2[2] * zs[k]) / c1;
    }
    double x1 = lx[0], x2 = lx[1], x3 = lx[2], y1 = ly[0], y2 = ly[1], y3 = ly[2];
    double Area = 0.5 * fabs((x1 - x3) * (y2 - y1) - (x1 - x2) * (y3 - y1));
    cplx coeff = gamma / ((2 * Area) * (2 * Area));
    tri_block(lx, ly, edge_lengths, Area, coeff, Bmat);
}
