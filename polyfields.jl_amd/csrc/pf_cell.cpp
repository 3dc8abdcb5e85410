// Unit-cell geometry and contour discretisation (host, FP64).
//
// Restates src/system/basis.jl:11-116 (make_basis / basis_matrix),
// src/system/cell.jl:113-258 (basis!, dbasis!, ddbasis!) and
// src/species/base.jl:27-56 (chain_grid).  Matrices are row-major 3x3 arrays with
// R[i*3+j] = R_ij; column j of R is lattice vector a_j as in the reference's SMat3D.
#include "pf_host.h"

#include <cmath>
#include <cstring>

namespace pf {

static bool isapprox(double a, double b) {
    const double rtol = std::sqrt(2.220446049250313e-16);
    return a == b || std::fabs(a - b) <= rtol * std::max(std::fabs(a), std::fabs(b));
}

static void basis_matrix(double a, double b, double c, double alpha, double beta, double gamma,
                         double* R) {
    const double p2 = M_PI / 2;
    for (int i = 0; i < 9; ++i) R[i] = 0.0;
    if (isapprox(alpha, p2) && isapprox(beta, p2) && isapprox(gamma, p2)) {
        R[0] = a;
        R[4] = b;
        R[8] = c;
        return;
    }
    const double ca = isapprox(alpha, p2) ? 0.0 : std::cos(alpha);
    const double cb = isapprox(beta, p2) ? 0.0 : std::cos(beta);
    double cg, sg;
    if (isapprox(gamma, p2)) {
        cg = 0.0;
        sg = 1.0;
    } else {
        cg = std::cos(gamma);
        sg = std::sin(gamma);
    }
    const double bx = b * cg, by = b * sg;
    const double cx = c * cb, cy = c * (ca - cb * cg) / sg;
    const double cz = std::sqrt(c * c - cx * cx - cy * cy);
    R[0] = a;  R[1] = bx; R[2] = cx;
    R[3] = 0;  R[4] = by; R[5] = cy;
    R[6] = 0;  R[7] = 0;  R[8] = cz;
}

int crystal_nparams(int dim, const std::string& cr) {
    if (dim == 1 && cr == "lamellar") return 1;
    if (dim == 2) {
        if (cr == "square" || cr == "hexagonal") return 1;
        if (cr == "rectangular") return 2;
        if (cr == "oblique") return 3;
    }
    if (dim == 3) {
        if (cr == "cubic" || cr == "pnna" || cr == "fddd1") return 1;
        if (cr == "tetragonal" || cr == "hexagonal") return 2;
        if (cr == "orthorhombic") return 3;
        if (cr == "monoclinic") return 4;
        if (cr == "triclinic") return 6;
    }
    return -1;
}

void Cell::basis() {
    const double pi = M_PI;
    const double* p = params.data();
    const int np = crystal_nparams(dim, crystal);
    if (np < 0)
        throw Error(-2, "Unrecognized crystal system `" + crystal + "` for dim = " +
                            std::to_string(dim) + ".");
    if ((int)params.size() != np)
        throw Error(-2, "crystal `" + crystal + "` takes " + std::to_string(np) + " parameters");
    if (dim == 1) {
        basis_matrix(p[0], 1.0, 1.0, pi / 2, pi / 2, pi / 2, R);
    } else if (dim == 2) {
        if (crystal == "square") basis_matrix(p[0], p[0], 1.0, pi / 2, pi / 2, pi / 2, R);
        else if (crystal == "rectangular") basis_matrix(p[0], p[1], 1.0, pi / 2, pi / 2, pi / 2, R);
        else if (crystal == "hexagonal") basis_matrix(p[0], p[0], 1.0, pi / 2, pi / 2, pi / 3, R);
        else basis_matrix(p[0], p[1], 1.0, pi / 2, pi / 2, p[2], R);
    } else {
        if (crystal == "cubic") basis_matrix(p[0], p[0], p[0], pi / 2, pi / 2, pi / 2, R);
        else if (crystal == "tetragonal") basis_matrix(p[0], p[0], p[1], pi / 2, pi / 2, pi / 2, R);
        else if (crystal == "orthorhombic") basis_matrix(p[0], p[1], p[2], pi / 2, pi / 2, pi / 2, R);
        else if (crystal == "hexagonal") basis_matrix(p[0], p[0], p[1], pi / 2, pi / 2, pi / 3, R);
        else if (crystal == "monoclinic") basis_matrix(p[0], p[1], p[2], pi / 2, p[3], pi / 2, R);
        else if (crystal == "triclinic") basis_matrix(p[0], p[1], p[2], p[3], p[4], p[5], R);
        else if (crystal == "pnna")
            basis_matrix(2.0 * p[0], std::sqrt(3.0) * p[0], p[0], pi / 2, pi / 2, pi / 2, R);
        else
            basis_matrix(p[0], std::sqrt(3.0) * p[0], 2.0 * std::sqrt(3.0) * p[0], pi / 2, pi / 2,
                         pi / 2, R);
    }
    // G = 2 pi inv(R)^T ; volume = det(R)
    const double a = R[0], b = R[1], c = R[2], d = R[3], e = R[4], f = R[5], g = R[6], h = R[7],
                 k = R[8];
    const double A = e * k - f * h, B = -(d * k - f * g), C = d * h - e * g;
    const double D = -(b * k - c * h), E = a * k - c * g, F = -(a * h - b * g);
    const double Gc = b * f - c * e, H = -(a * f - c * d), I = a * e - b * d;
    volume = a * A + b * B + c * C;
    // inv(R) = adj / det, adj^T = cofactor matrix; inv(R)^T = cofactor / det
    const double cof[9] = {A, B, C, D, E, F, Gc, H, I};
    for (int i = 0; i < 9; ++i) G[i] = 2 * M_PI * cof[i] / volume;
}

void Cell::dbasis() {
    const int np = (int)params.size();
    dR.assign(9 * np, 0.0);
    auto set = [&](int i, int j, int k, double v) { dR[(i * 3 + j) * np + k] = v; };
    const double* p = params.data();
    if (dim == 1) {
        set(0, 0, 0, 1.0);
    } else if (dim == 2) {
        if (crystal == "square") {
            set(0, 0, 0, 1.0); set(1, 1, 0, 1.0);
        } else if (crystal == "rectangular") {
            set(0, 0, 0, 1.0); set(1, 1, 1, 1.0);
        } else if (crystal == "hexagonal") {
            set(0, 0, 0, 1.0); set(0, 1, 0, 0.5); set(1, 1, 0, std::sqrt(0.75));
        } else {
            const double b = p[1], gm = p[2];
            set(0, 0, 0, 1.0);
            set(0, 1, 1, std::cos(gm)); set(1, 1, 1, std::sin(gm));
            set(0, 1, 2, -b * std::sin(gm)); set(1, 1, 2, b * std::cos(gm));
        }
    } else {
        if (crystal == "cubic") {
            set(0, 0, 0, 1.0); set(1, 1, 0, 1.0); set(2, 2, 0, 1.0);
        } else if (crystal == "tetragonal") {
            set(0, 0, 0, 1.0); set(1, 1, 0, 1.0); set(2, 2, 1, 1.0);
        } else if (crystal == "orthorhombic") {
            set(0, 0, 0, 1.0); set(1, 1, 1, 1.0); set(2, 2, 2, 1.0);
        } else if (crystal == "hexagonal") {
            set(0, 0, 0, 1.0); set(0, 1, 0, 0.5); set(1, 1, 0, std::sqrt(0.75)); set(2, 2, 1, 1.0);
        } else if (crystal == "monoclinic") {
            const double c = p[2], be = p[3];
            set(0, 0, 0, 1.0); set(1, 1, 1, 1.0);
            set(0, 2, 2, std::cos(be)); set(2, 2, 2, std::sin(be));
            set(0, 2, 3, -c * std::sin(be)); set(2, 2, 3, c * std::cos(be));
        } else if (crystal == "triclinic") {
            const double b = p[1], c = p[2], al = p[3], be = p[4], gm = p[5];
            const double cy = R[1 * 3 + 2], cz = R[2 * 3 + 2];
            const double sa = std::sin(al), ca = std::cos(al), sb = std::sin(be), cb = std::cos(be);
            const double sg = std::sin(gm), cg = std::cos(gm), tg = std::tan(gm);
            set(0, 0, 0, 1.0);
            set(0, 1, 1, cg); set(1, 1, 1, sg);
            set(0, 2, 2, cb); set(1, 2, 2, cy / c); set(2, 2, 2, cz / c);
            set(1, 2, 3, -c * sa / sg);
            set(2, 2, 3, c * c * (ca - cb * cg) * sa / (sg * sg) / cz);
            set(0, 2, 4, -c * sb);
            set(1, 2, 4, c * sb / tg);
            set(2, 2, 4, c * c * (-ca / tg + cb / sg) * sb / sg / cz);
            set(0, 1, 5, -b * sg); set(1, 1, 5, b * cg);
            set(1, 2, 5, c * (-ca / tg + cb / sg) / sg);
            set(2, 2, 5, c * c * ((2 + std::cos(2 * al) + std::cos(2 * be)) / tg -
                                  ca * cb * (3 + std::cos(2 * gm)) / sg) /
                             (sg * sg) / cz / 2);
        } else if (crystal == "pnna") {
            set(0, 0, 0, 2.0); set(1, 1, 0, std::sqrt(3.0)); set(2, 2, 0, 1.0);
        } else if (crystal == "fddd1") {
            set(0, 0, 0, 1.0); set(1, 1, 0, std::sqrt(3.0)); set(2, 2, 0, 2.0 * std::sqrt(3.0));
        }
    }
    // dG_ijk = -(1/2pi) sum_lm G_il dR_mlk G_mj  (src/system/cell.jl:215-228)
    dG.assign(9 * np, 0.0);
    for (int k = 0; k < np; ++k)
        for (int i = 0; i < dim; ++i)
            for (int j = 0; j < dim; ++j)
                for (int l = 0; l < dim; ++l)
                    for (int m = 0; m < dim; ++m)
                        dG[(i * 3 + j) * np + k] -=
                            G[i * 3 + l] * dR[(m * 3 + l) * np + k] * G[m * 3 + j];
    for (auto& v : dG) v /= 2 * M_PI;
}

void Cell::ddbasis() {
    const int np = (int)params.size();
    dRR.assign(9 * np, 0.0);
    dGG.assign(9 * np, 0.0);
    for (int k = 0; k < np; ++k)
        for (int i = 0; i < dim; ++i)
            for (int j = 0; j < dim; ++j)
                for (int l = 0; l < dim; ++l) {
                    // Rt[i,l] = R[l,i]
                    dRR[(i * 3 + j) * np + k] += R[l * 3 + i] * dR[(l * 3 + j) * np + k] +
                                                 R[l * 3 + j] * dR[(l * 3 + i) * np + k];
                    dGG[(i * 3 + j) * np + k] += G[l * 3 + i] * dG[(l * 3 + j) * np + k] +
                                                 G[l * 3 + j] * dG[(l * 3 + i) * np + k];
                }
}

void Cell::update() {
    basis();
    dbasis();
    ddbasis();
    ++version;
}

void chain_grid(const std::vector<double>& N_block, int Ns, double* N_out,
                std::vector<int>& begin, std::vector<double>& ds_block) {
    double N = 0.0;
    for (double x : N_block) N += x;
    const int nb = (int)N_block.size();
    const double ds = N / Ns;
    begin.assign(nb + 1, 0);
    ds_block.assign(nb, 0.0);
    int bgn = 1;
    begin[0] = bgn;
    for (int i = 0; i < nb; ++i) {
        int half = (int)std::floor(N_block[i] / ds / 2.0 + 0.5);
        if (half == 0) {
            half = 1;
            ds_block[i] = N_block[i] / 2.0;
        } else {
            ds_block[i] = N_block[i] / half / 2.0;
        }
        bgn = bgn + half * 2;
        begin[i + 1] = bgn;
    }
    *N_out = N;
}

// ---- small dense linear algebra shared by the Anderson and cell updaters ----------
// LU with partial pivoting (first maximal pivot), inverse by substitution against the
// identity, then a row-by-row product: the same arithmetic order as the CPU oracle.
bool lu_inverse(int n, const double* U, double* inv) {
    // LU with partial pivoting, then the inverse column by column (inv!(lu!(U)),
    // updaters/anderson.jl:185).  The substitutions run row by row over all n right-hand
    // sides at once: every element sees the same operations in the same order as one
    // column at a time, and the independent columns vectorise.  Working storage is on the
    // stack for the sizes used here (Anderson histories up to PF_MAXHIST, cell Jacobians),
    // on the heap beyond.
    constexpr int kSmall = 16;
    double a_s[kSmall * kSmall], x_s[kSmall * kSmall];
    int perm_s[kSmall];
    std::vector<double> a_h, x_h;
    std::vector<int> perm_h;
    double* A = a_s;
    double* X = x_s;
    int* perm = perm_s;
    if (n > kSmall) {
        a_h.resize((size_t)n * n);
        x_h.resize((size_t)n * n);
        perm_h.resize(n);
        A = a_h.data();
        X = x_h.data();
        perm = perm_h.data();
    }
    for (int i = 0; i < n * n; ++i) A[i] = U[i];
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = std::fabs(A[k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(A[i * n + k]) > best) {
                best = std::fabs(A[i * n + k]);
                p = i;
            }
        if (A[p * n + k] == 0.0) return false;
        if (p != k) {
            for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[p * n + j]);
            std::swap(perm[k], perm[p]);
        }
        for (int i = k + 1; i < n; ++i) {
            A[i * n + k] = A[i * n + k] / A[k * n + k];
            const double f = A[i * n + k];
            for (int j = k + 1; j < n; ++j) A[i * n + j] = A[i * n + j] - f * A[k * n + j];
        }
    }
    // forward substitution: row i of L^-1 P (X holds Y, row-major [i][c])
    for (int i = 0; i < n; ++i) {
        double* xi = X + (size_t)i * n;
        for (int c = 0; c < n; ++c) xi[c] = (perm[i] == c) ? 1.0 : 0.0;
        for (int j = 0; j < i; ++j) {
            const double f = A[i * n + j];
            const double* xj = X + (size_t)j * n;
            for (int c = 0; c < n; ++c) xi[c] -= f * xj[c];
        }
    }
    // back substitution in place: rows from the last up
    for (int i = n - 1; i >= 0; --i) {
        double* xi = X + (size_t)i * n;
        for (int j = i + 1; j < n; ++j) {
            const double f = A[i * n + j];
            const double* xj = X + (size_t)j * n;
            for (int c = 0; c < n; ++c) xi[c] -= f * xj[c];
        }
        const double d = A[i * n + i];
        for (int c = 0; c < n; ++c) xi[c] = xi[c] / d;
    }
    for (int i = 0; i < n * n; ++i) inv[i] = X[i];
    return true;
}

void matvec(int n, const double* M, const double* v, double* out) {
    for (int i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int j = 0; j < n; ++j) acc += M[i * n + j] * v[j];
        out[i] = acc;
    }
}

}  // namespace pf
