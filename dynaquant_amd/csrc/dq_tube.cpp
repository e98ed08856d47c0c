// dq_tube.cpp -- host generator of the cylindrical dipole aggregates used by
// the benchmark systems.  Pairs are searched ring-window by ring-window
// (|dz| < cutoff), each J_ij is computed once for i < j and mirrored, so the
// CSR is exactly symmetric.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/dynaquant.h"

namespace {
thread_local std::string g_tube_err;

struct Site { double p[3], u[3]; int wall, ring; };

bool make_sites(const dq_tube_spec* s, std::vector<Site>& out)
{
    if (!s || s->n_walls < 1 || s->n_walls > 4 || s->per_ring < 1) return false;
    int n = 0;
    for (int w = 0; w < s->n_walls; ++w) {
        if (s->rings[w] < 1) return false;
        n += s->rings[w] * s->per_ring;
    }
    if (s->order == 1)
        for (int w = 1; w < s->n_walls; ++w)
            if (s->rings[w] != s->rings[0]) return false;
    out.assign(n, Site{});
    const double pi = std::atan(1.0) * 4.0;
    const double tilt = s->tilt_deg * pi / 180.0;
    const double ct = std::cos(tilt), st = std::sin(tilt);
    int off = 0;
    for (int w = 0; w < s->n_walls; ++w) {
        for (int r = 0; r < s->rings[w]; ++r)
            for (int k = 0; k < s->per_ring; ++k) {
                const int i = s->order == 0 ? off + r * s->per_ring + k
                                            : r * (s->n_walls * s->per_ring) + w * s->per_ring + k;
                double phi = 2.0 * pi * (double)k / (double)s->per_ring;
                if (r & 1) phi += pi / (double)s->per_ring;
                const double cp = std::cos(phi), sp = std::sin(phi);
                Site& t = out[i];
                t.p[0] = s->radius_nm[w] * cp;
                t.p[1] = s->radius_nm[w] * sp;
                t.p[2] = s->ring_dz_nm * (double)r;
                t.u[0] = s->mu_debye * (ct * -sp);
                t.u[1] = s->mu_debye * (ct * cp);
                t.u[2] = s->mu_debye * st;
                t.wall = w;
                t.ring = r;
            }
        off += s->rings[w] * s->per_ring;
    }
    return true;
}

double coupling(const Site& a, const Site& b, double kappa, double* r_out)
{
    const double dx = b.p[0] - a.p[0], dy = b.p[1] - a.p[1], dz = b.p[2] - a.p[2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double r = std::sqrt(r2);
    *r_out = r;
    const double ux = dx / r, uy = dy / r, uz = dz / r;
    const double mm = a.u[0] * b.u[0] + a.u[1] * b.u[1] + a.u[2] * b.u[2];
    const double ar = a.u[0] * ux + a.u[1] * uy + a.u[2] * uz;
    const double br = b.u[0] * ux + b.u[1] * uy + b.u[2] * uz;
    return kappa * (mm - 3.0 * ar * br) / (r2 * r);
}
}  // namespace

extern "C" dq_status dq_tube_sites(const dq_tube_spec* spec, double* pos, double* dip, int32_t* wall,
                                   int32_t* n_out)
{
    std::vector<Site> s;
    if (!make_sites(spec, s)) return DQ_ERR_INVALID_ARG;
    for (size_t i = 0; i < s.size(); ++i) {
        for (int c = 0; c < 3; ++c) {
            if (pos) pos[3 * i + c] = s[i].p[c];
            if (dip) dip[3 * i + c] = s[i].u[c];
        }
        if (wall) wall[i] = s[i].wall;
    }
    if (n_out) *n_out = (int32_t)s.size();
    return DQ_OK;
}

extern "C" dq_status dq_tube_couplings(const dq_tube_spec* spec, int32_t* rowptr, int32_t* col, double* val,
                                       int64_t* nnz_out)
{
    std::vector<Site> s;
    if (!make_sites(spec, s) || !rowptr) return DQ_ERR_INVALID_ARG;
    const int n = (int)s.size();
    const double cut = spec->cutoff_nm;
    // sites sorted by z (ring index); window over |dz| < cutoff
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return s[a].p[2] < s[b].p[2]; });
    std::vector<std::vector<std::pair<int, double>>> rows(n);
    for (int a = 0; a < n; ++a) {
        const int i = order[a];
        for (int c = a + 1; c < n; ++c) {
            const int j = order[c];
            if (s[j].p[2] - s[i].p[2] >= cut) break;
            double r;
            const int lo = std::min(i, j), hi = std::max(i, j);
            const double J = coupling(s[lo], s[hi], spec->kappa, &r);
            if (!(r < cut)) continue;
            rows[lo].push_back({hi, J});
            rows[hi].push_back({lo, J});
        }
    }
    int64_t nnz = 0;
    rowptr[0] = 0;
    for (int i = 0; i < n; ++i) {
        std::sort(rows[i].begin(), rows[i].end());
        if (col && val)
            for (size_t k = 0; k < rows[i].size(); ++k) {
                col[nnz + k] = rows[i][k].first;
                val[nnz + k] = rows[i][k].second;
            }
        nnz += (int64_t)rows[i].size();
        rowptr[i + 1] = (int32_t)nnz;
    }
    if (nnz_out) *nnz_out = nnz;
    return DQ_OK;
}
