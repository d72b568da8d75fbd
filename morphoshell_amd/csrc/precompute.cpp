// Host set-up of the per-step path: mesh info, patch fitting, the patch
// derivative operators G and H, node masses and the characteristic scales.
//
// These are the reference's once-per-run functions (calc_mesh_info,
// set_up_patch_fitting, calc_deriv_mats, calc_dof_masses, calc_*_scales,
// calc_timestep, ...).  Their outputs are the per-triangle records the
// device kernels stream every step, so their conventions (slot order, sign
// and scaling of G and H, the patch-choice rule) are the kernels' input
// format.  Written for this project from the mathematics: G interpolates a
// linear function through the three vertices, H extracts the second
// derivatives of the quadratic through the six patch nodes.
//
// Per-triangle work is independent and runs under OpenMP.

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "ms_common.h"

namespace ms {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }
const std::string& last_error() { return g_last_error; }

namespace {

// Vertex -> incident triangles, ascending triangle id
// (calc_triangles_incident_on_nodes.cpp:42-53).
struct Incidence {
    std::vector<int64_t> off;
    std::vector<int32_t> tri;
};

Incidence vertex_incidence(int32_t N, int32_t T, const int32_t* tv) {
    Incidence inc;
    inc.off.assign((size_t)N + 1, 0);
    for (int64_t i = 0; i < 3LL * T; ++i) inc.off[(size_t)tv[i] + 1]++;
    for (int32_t n = 0; n < N; ++n) inc.off[n + 1] += inc.off[n];
    inc.tri.resize((size_t)3 * T);
    std::vector<int64_t> cur(inc.off.begin(), inc.off.end() - 1);
    for (int32_t t = 0; t < T; ++t)
        for (int v = 0; v < 3; ++v) {
            int32_t n = tv[3 * (int64_t)t + v];
            // a degenerate triangle listing a node twice is rejected earlier
            inc.tri[(size_t)cur[n]++] = t;
        }
    return inc;
}

bool check_ids(int32_t N, int32_t T, const int32_t* tv, std::string& err) {
    bool zero = false;
    for (int64_t i = 0; i < 3LL * T; ++i) {
        if (tv[i] < 0 || tv[i] >= N) {
            err = fmt("triangle %lld has vertex id %d outside [0, %d)", (long long)(i / 3), tv[i], N);
            return false;
        }
        if (tv[i] == 0) zero = true;
    }
    for (int32_t t = 0; t < T; ++t) {
        const int32_t* v = tv + 3 * (int64_t)t;
        if (v[0] == v[1] || v[1] == v[2] || v[0] == v[2]) {
            err = fmt("triangle %d repeats a vertex", t);
            return false;
        }
    }
    if (!zero) {  // read_vtk_data.cpp:163-175
        err = "node ids must start from 0 (no triangle uses node 0)";
        return false;
    }
    return true;
}

// Full-pivot LU of a 6x6 matrix; returns false if numerically singular with
// the rank rule of a full-pivot LU (pivot <= 6*eps*max pivot counts as zero).
// On success fills inv with the inverse.
bool invert6(const double A_in[6][6], double inv[6][6]) {
    double A[6][6];
    std::memcpy(A, A_in, sizeof(A));
    int rp[6], cp[6];
    for (int i = 0; i < 6; ++i) rp[i] = cp[i] = i;
    double maxpivot = 0.0;
    for (int k = 0; k < 6; ++k) {
        int bi = k, bj = k;
        double best = -1.0;
        for (int i = k; i < 6; ++i)
            for (int j = k; j < 6; ++j)
                if (std::fabs(A[i][j]) > best) {
                    best = std::fabs(A[i][j]);
                    bi = i;
                    bj = j;
                }
        if (k == 0) maxpivot = best;
        if (!(best > 6.0 * 2.220446049250313e-16 * maxpivot) || best == 0.0) return false;
        if (bi != k) {
            for (int j = 0; j < 6; ++j) std::swap(A[k][j], A[bi][j]);
            std::swap(rp[k], rp[bi]);
        }
        if (bj != k) {
            for (int i = 0; i < 6; ++i) std::swap(A[i][k], A[i][bj]);
            std::swap(cp[k], cp[bj]);
        }
        for (int i = k + 1; i < 6; ++i) {
            double l = A[i][k] / A[k][k];
            A[i][k] = l;
            for (int j = k + 1; j < 6; ++j) A[i][j] -= l * A[k][j];
        }
    }
    // Solve (P A Q) y = P e_c for each unit vector; inverse = Q y P.
    for (int c = 0; c < 6; ++c) {
        double y[6];
        for (int i = 0; i < 6; ++i) y[i] = (rp[i] == c) ? 1.0 : 0.0;
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < i; ++j) y[i] -= A[i][j] * y[j];
        for (int i = 5; i >= 0; --i) {
            for (int j = i + 1; j < 6; ++j) y[i] -= A[i][j] * y[j];
            y[i] /= A[i][i];
        }
        for (int i = 0; i < 6; ++i) inv[cp[i]][c] = y[i];
    }
    return true;
}

// Largest singular value of a 6x3 matrix: sqrt of the largest eigenvalue of
// the 3x3 Gram matrix, by cyclic Jacobi rotations.
double sigma_max_6x3(const double H[6][3]) {
    double S[3][3] = {};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int n = 0; n < 6; ++n) S[i][j] += H[n][i] * H[n][j];
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = S[0][1] * S[0][1] + S[0][2] * S[0][2] + S[1][2] * S[1][2];
        double diag = S[0][0] * S[0][0] + S[1][1] * S[1][1] + S[2][2] * S[2][2];
        if (off <= 1e-32 * diag) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (S[p][q] == 0.0) continue;
                double theta = (S[q][q] - S[p][p]) / (2.0 * S[p][q]);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) {  // S <- J^T S J
                    double skp = S[k][p], skq = S[k][q];
                    S[k][p] = c * skp - s * skq;
                    S[k][q] = s * skp + c * skq;
                }
                for (int k = 0; k < 3; ++k) {
                    double spk = S[p][k], sqk = S[q][k];
                    S[p][k] = c * spk - s * sqk;
                    S[q][k] = s * spk + c * sqk;
                }
            }
    }
    double lmax = std::max(S[0][0], std::max(S[1][1], S[2][2]));
    return std::sqrt(std::max(lmax, 0.0));
}

}  // namespace
}  // namespace ms

using namespace ms;

extern "C" const char* ms_last_error(void) { return ms::last_error().c_str(); }

extern "C" int ms_mesh_info(int32_t N, int32_t T, const double* pos, const int32_t* tv, double* areas,
                            double* centroids, uint8_t* boundary_nodes, ms_mesh_info_out* out) {
    if (N <= 0 || T <= 0 || !pos || !tv || !out) {
        set_last_error("ms_mesh_info: bad arguments");
        return MS_ERR_ARG;
    }
    std::string err;
    if (!check_ids(N, T, tv, err)) {
        set_last_error("ms_mesh_info: " + err);
        return MS_ERR_MESH;
    }
    const double* X = pos;
    const double* Y = pos + N;
    const double* Z = pos + 2 * (int64_t)N;
    Incidence inc = vertex_incidence(N, T, tv);
    for (int32_t n = 0; n < N; ++n)
        if (inc.off[n + 1] == inc.off[n]) {
            set_last_error(fmt("ms_mesh_info: node %d touches no triangle", n));
            return MS_ERR_MESH;
        }
    // Edges (a < b) from each triangle, counted by the number of triangles
    // sharing them (calc_edges.cpp: shared by 1 = boundary, 2 = interior).
    std::vector<std::pair<int64_t, int32_t>> e;  // key, tri
    e.reserve((size_t)3 * T);
    for (int32_t t = 0; t < T; ++t)
        for (int v = 0; v < 3; ++v) {
            int64_t a = tv[3 * (int64_t)t + v], b = tv[3 * (int64_t)t + (v + 1) % 3];
            if (a > b) std::swap(a, b);
            e.push_back({a * (int64_t)N + b, t});
        }
    std::sort(e.begin(), e.end());
    std::vector<double> bd_len;
    std::vector<uint8_t> bnode((size_t)N, 0);
    int64_t nedges = 0, nbedges = 0;
    for (size_t i = 0; i < e.size();) {
        size_t j = i;
        while (j < e.size() && e[j].first == e[i].first) ++j;
        size_t cnt = j - i;
        if (cnt > 2) {
            set_last_error(fmt("ms_mesh_info: an edge is shared by %zu triangles", cnt));
            return MS_ERR_MESH;
        }
        ++nedges;
        if (cnt == 1) {
            ++nbedges;
            int64_t a = e[i].first / N, b = e[i].first % N;
            bnode[a] = bnode[b] = 1;
            double dx = X[b] - X[a], dy = Y[b] - Y[a], dz = Z[b] - Z[a];
            bd_len.push_back(std::sqrt(dx * dx + dy * dy + dz * dz));
        }
        i = j;
    }
    // 3T = 2E - E_b and a non-negative hole count (calc_mesh_info.cpp:122-131)
    if (3LL * T != 2 * nedges - nbedges) {
        set_last_error("ms_mesh_info: edge and triangle counts violate 3T = 2E - E_b");
        return MS_ERR_MESH;
    }
    int64_t holes = 1 + nedges - N - (int64_t)T;
    if (holes < 0) {
        set_last_error("ms_mesh_info: negative hole count (does every node touch a triangle?)");
        return MS_ERR_MESH;
    }
    std::vector<double> area((size_t)T);
    double min_alt = INFINITY;
#pragma omp parallel for reduction(min : min_alt) schedule(static)
    for (int32_t t = 0; t < T; ++t) {
        const int32_t* v = tv + 3 * (int64_t)t;
        double s1[3] = {X[v[1]] - X[v[0]], Y[v[1]] - Y[v[0]], Z[v[1]] - Z[v[0]]};
        double s2[3] = {X[v[2]] - X[v[0]], Y[v[2]] - Y[v[0]], Z[v[2]] - Z[v[0]]};
        double s3[3] = {X[v[2]] - X[v[1]], Y[v[2]] - Y[v[1]], Z[v[2]] - Z[v[1]]};
        double c[3] = {0.5 * (s1[1] * s2[2] - s1[2] * s2[1]), 0.5 * (s1[2] * s2[0] - s1[0] * s2[2]),
                       0.5 * (s1[0] * s2[1] - s1[1] * s2[0])};
        double A = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        area[t] = A;
        double l1 = std::sqrt(s1[0] * s1[0] + s1[1] * s1[1] + s1[2] * s1[2]);
        double l2 = std::sqrt(s2[0] * s2[0] + s2[1] * s2[1] + s2[2] * s2[2]);
        double l3 = std::sqrt(s3[0] * s3[0] + s3[1] * s3[1] + s3[2] * s3[2]);
        double alt = std::min(2.0 * A / l1, std::min(2.0 * A / l2, 2.0 * A / l3));
        min_alt = std::min(min_alt, alt);
        if (centroids) {
            centroids[t] = (X[v[0]] + X[v[1]] + X[v[2]]) / 3.0;
            centroids[T + t] = (Y[v[0]] + Y[v[1]] + Y[v[2]]) / 3.0;
            centroids[2 * (int64_t)T + t] = (Z[v[0]] + Z[v[1]] + Z[v[2]]) / 3.0;
        }
    }
    if (areas) std::memcpy(areas, area.data(), sizeof(double) * (size_t)T);
    if (boundary_nodes) std::memcpy(boundary_nodes, bnode.data(), (size_t)N);
    out->perimeter = kahan_sum(bd_len.begin(), bd_len.end());
    out->total_ref_area = kahan_sum(area.begin(), area.end());
    out->approx_min_tri_size = min_alt;
    out->num_edges = (int32_t)nedges;
    out->num_boundary_edges = (int32_t)nbedges;
    int32_t nb = 0;
    for (auto b : bnode) nb += b;
    out->num_boundary_nodes = nb;
    out->num_holes = (int32_t)holes;
    return MS_OK;
}

extern "C" int ms_patch_fit(int32_t N, int32_t T, const double* pos, const int32_t* tv, double thresh,
                            int32_t* patch_ids, double* Gout, double* Hout, int32_t* num_retried) {
    if (N <= 0 || T <= 0 || !pos || !tv || !patch_ids || !Gout || !Hout || !(thresh > 0)) {
        set_last_error("ms_patch_fit: bad arguments");
        return MS_ERR_ARG;
    }
    std::string err;
    if (!check_ids(N, T, tv, err)) {
        set_last_error("ms_patch_fit: " + err);
        return MS_ERR_MESH;
    }
    const double* X = pos;
    const double* Y = pos + N;
    const double* Z = pos + 2 * (int64_t)N;
    Incidence inc = vertex_incidence(N, T, tv);
    int32_t failed_tri = -1;
    int64_t retried = 0;

#pragma omp parallel reduction(+ : retried)
    {
        std::vector<int32_t> cand;
        std::vector<double> dist;
        std::vector<int32_t> order;
#pragma omp for schedule(dynamic, 256)
        for (int32_t t = 0; t < T; ++t) {
            const int32_t* v = tv + 3 * (int64_t)t;
            const double cx = (X[v[0]] + X[v[1]] + X[v[2]]) / 3.0;
            const double cy = (Y[v[0]] + Y[v[1]] + Y[v[2]]) / 3.0;
            const double cz = (Z[v[0]] + Z[v[1]] + Z[v[2]]) / 3.0;
            // Candidates: vertices of the triangles incident on t's vertices,
            // in first-seen order, excluding t's own vertices
            // (set_up_patch_fitting.cpp:73-114).
            cand.clear();
            dist.clear();
            for (int vv = 0; vv < 3; ++vv) {
                int32_t n = v[vv];
                for (int64_t k = inc.off[n]; k < inc.off[n + 1]; ++k) {
                    int32_t s = inc.tri[(size_t)k];
                    for (int w = 0; w < 3; ++w) {
                        int32_t m = tv[3 * (int64_t)s + w];
                        if (m == v[0] || m == v[1] || m == v[2]) continue;
                        if (std::find(cand.begin(), cand.end(), m) != cand.end()) continue;
                        double dx = X[m] - cx, dy = Y[m] - cy, dz = Z[m] - cz;
                        cand.push_back(m);
                        dist.push_back(std::sqrt(dx * dx + dy * dy + dz * dz));
                    }
                }
            }
            // Ascending distance; ties keep first-seen order (the reference's
            // std::sort leaves tie order unspecified, set_up_patch_fitting.cpp:125).
            order.resize(cand.size());
            std::iota(order.begin(), order.end(), 0);
            std::stable_sort(order.begin(), order.end(),
                             [&](int32_t a, int32_t b) { return dist[(size_t)a] < dist[(size_t)b]; });
            bool ok = false;
            int32_t pn[3] = {-1, -1, -1};
            double Hbest[6][3];
            const size_t nc = cand.size();
            if (nc >= 3) {
                pn[0] = cand[(size_t)order[0]];
                bool first = true;
                for (size_t p = 1; p < std::min<size_t>(4, nc) && !ok; ++p) {
                    pn[1] = cand[(size_t)order[p]];
                    for (size_t q = 2; q < nc; ++q) {
                        pn[2] = cand[(size_t)order[q]];
                        const int32_t ids[6] = {v[0], v[1], v[2], pn[0], pn[1], pn[2]};
                        double size2 = 0.0;
                        for (int n = 0; n < 6; ++n) {
                            double dx = X[ids[n]] - cx, dy = Y[ids[n]] - cy, dz = Z[ids[n]] - cz;
                            size2 += dx * dx + dy * dy + dz * dz;
                        }
                        size2 /= 6.0;  // squared RMS patch size
                        bool good = (pn[1] != pn[2]);
                        double Hc[6][3];
                        if (good) {
                            // Column n of the fit matrix holds the quadratic
                            // basis (1, dx, dy, dx^2/2, dx dy, dy^2/2) at node n.
                            double M6[6][6], inv[6][6];
                            for (int n = 0; n < 6; ++n) {
                                double dx = X[ids[n]] - cx, dy = Y[ids[n]] - cy;
                                M6[0][n] = 1.0;
                                M6[1][n] = dx;
                                M6[2][n] = dy;
                                M6[3][n] = 0.5 * dx * dx;
                                M6[4][n] = dx * dy;
                                M6[5][n] = 0.5 * dy * dy;
                            }
                            good = invert6(M6, inv);
                            if (good) {
                                for (int n = 0; n < 6; ++n)
                                    for (int k = 0; k < 3; ++k) Hc[n][k] = inv[n][3 + k];
                                double ratio = sigma_max_6x3(Hc) * size2 / thresh;
                                good = (ratio < 1.0);
                            }
                        }
                        if (!good) {
                            if (first) ++retried;
                            first = false;
                            continue;
                        }
                        std::memcpy(Hbest, Hc, sizeof(Hbest));
                        ok = true;
                        break;
                    }
                    first = false;
                }
            }
            if (!ok) {
#pragma omp critical
                {
                    if (failed_tri < 0 || t < failed_tri) failed_tri = t;
                }
                continue;
            }
            int32_t* pid = patch_ids + 6 * (int64_t)t;
            pid[0] = v[0];
            pid[1] = v[1];
            pid[2] = v[2];
            pid[3] = pn[0];
            pid[4] = pn[1];
            pid[5] = pn[2];
            double* H = Hout + 18 * (int64_t)t;
            for (int n = 0; n < 6; ++n)
                for (int k = 0; k < 3; ++k) H[3 * n + k] = Hbest[n][k];
            // G: rows 1,2 of inv([1, dx_v, dy_v]) (calc_deriv_mats.cpp:70-91),
            // by the explicit 3x3 inverse.
            double dx[3], dy[3];
            for (int i = 0; i < 3; ++i) {
                dx[i] = X[v[i]] - cx;
                dy[i] = Y[v[i]] - cy;
            }
            // det of [[1,dx0,dy0],[1,dx1,dy1],[1,dx2,dy2]]
            double det = (dx[1] * dy[2] - dx[2] * dy[1]) - (dx[0] * dy[2] - dx[2] * dy[0]) +
                         (dx[0] * dy[1] - dx[1] * dy[0]);
            double* G = Gout + 6 * (int64_t)t;
            // inverse row 1 (d/dx): cofactors of column 1
            G[0] = (dy[1] - dy[2]) / det;
            G[1] = (dy[2] - dy[0]) / det;
            G[2] = (dy[0] - dy[1]) / det;
            // inverse row 2 (d/dy)
            G[3] = (dx[2] - dx[1]) / det;
            G[4] = (dx[0] - dx[2]) / det;
            G[5] = (dx[1] - dx[0]) / det;
        }
    }
    if (num_retried) *num_retried = (int32_t)retried;
    if (failed_tri >= 0) {
        set_last_error(fmt("At least one search for patch nodes was exhausted without success (triangle %d); "
                           "all patch matrices in the search had a condition number above the acceptance "
                           "threshold. Try increasing patch_mat_dimless_conditioning_thresh.",
                           failed_tri));
        return MS_ERR_PATCH_SEARCH;
    }
    return MS_OK;
}

extern "C" int ms_dof_masses(int32_t N, int32_t T, const int32_t* tv, const double* areas, const double* h,
                             double rho, double* m) {
    if (N <= 0 || T <= 0 || !tv || !areas || !h || !m) {
        set_last_error("ms_dof_masses: bad arguments");
        return MS_ERR_ARG;
    }
    for (int32_t n = 0; n < N; ++n) m[n] = 0.0;
    // One third of each triangle's mass to each vertex, in triangle order
    // (calc_dof_masses.cpp:50-58).
    for (int32_t t = 0; t < T; ++t)
        for (int v = 0; v < 3; ++v) {
            int32_t n = tv[3 * (int64_t)t + v];
            if (n < 0 || n >= N) {
                set_last_error("ms_dof_masses: vertex id out of range");
                return MS_ERR_ARG;
            }
            m[n] += rho * areas[t] * h[t] / 3.0;
        }
    return MS_OK;
}

extern "C" int ms_calc_scales(int32_t T, const double* h, const double* mu, const double* areas,
                              double perimeter, double min_size, ms_scales* s) {
    if (T <= 0 || !h || !mu || !areas || !s || !(perimeter > 0) || !(min_size > 0)) {
        set_last_error("ms_calc_scales: bad arguments");
        return MS_ERR_ARG;
    }
    const double pi = 3.14159265358979323846;
    double hmin = INFINITY, mumin = INFINITY;
    for (int32_t t = 0; t < T; ++t) {
        hmin = std::min(hmin, h[t]);
        mumin = std::min(mumin, mu[t]);
    }
    const double nu = s->poisson_ratio, rho = s->ref_density;
    s->min_ref_thickness = hmin;
    s->min_ref_shear_modulus = mumin;
    // calc_char_mechanics_scales.cpp:59-74
    s->char_flexural_rigidity = (1.0 / 6.0) * mumin * hmin * hmin * hmin / (1.0 - nu);
    s->char_force_scale = mumin * min_size * hmin;
    s->char_speed_scale = std::sqrt(mumin / rho);
    s->char_stretch_energy_density_scale = mumin * hmin;
    s->char_stretch_energy_scale = s->char_stretch_energy_density_scale * kahan_sum(areas, areas + T);
    // calc_long_timescales.cpp:48-58
    const double k = 2.0 * pi / perimeter;
    const double D = s->char_flexural_rigidity;
    s->stretch_long_time = 2.0 * pi * std::sqrt(rho / mumin) / k;
    s->bend_long_time = 2.0 * pi * std::sqrt(rho * hmin / D) / (k * k);
    s->char_long_time = std::max(s->stretch_long_time, s->bend_long_time);
    // calc_damping_factor.cpp:50-56
    s->damping_factor = s->damping_prefactor * 2.0 * std::sqrt(D * rho * hmin * hmin * hmin) * k * k;
    // calc_timestep.cpp:52-71
    double st = 0.185 * s->timestep_prefactor * std::sqrt(rho / mumin) * min_size;
    double bt = s->timestep_prefactor * std::sqrt(rho * hmin / D) * min_size * min_size / (2 * pi);
    if (s->gradient_descent) {
        st = 0.185 * s->timestep_prefactor * (s->damping_factor / (mumin * hmin * hmin)) * min_size * min_size /
             (2.0 * pi);
        bt = s->timestep_prefactor * (s->damping_factor / (D * hmin)) * min_size * min_size * min_size * min_size /
             (2.0 * pi * 2.0 * pi * 2.0 * pi);
    }
    s->timestep = (st < bt) ? st : bt;
    // calc_dialling_procedure_params.cpp:68-77
    s->dial_phase_time = s->dial_phase_time_prefactor * s->char_long_time;
    s->timesteps_per_dial_phase = std::lround(s->dial_phase_time / s->timestep);
    s->timesteps_between_equil_checks =
        std::lround(s->time_between_equil_checks_prefactor * s->dial_phase_time / s->timestep);
    // calc_steps_between_outputs.cpp:48-69
    if (s->outputs_per_char_long_time < 0.0) {
        s->steps_between_outputs = -1;
    } else if (s->outputs_per_char_long_time < 1.0) {
        set_last_error("outputs_per_char_long_time must not lie in [0, 1)");
        return MS_ERR_ARG;
    } else {
        s->steps_between_outputs =
            std::lround(s->char_long_time / (s->timestep * s->outputs_per_char_long_time)) + 1;
    }
    if (s->gradient_descent) s->steps_between_outputs = 1000;
    return MS_OK;
}
