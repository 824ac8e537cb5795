// ecmc_common.cpp -- host-side helpers (see ecmc_common.h).
#include "ecmc_common.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

#include "../../include/jellyfysh_amd.h"

static thread_local char g_last_error[512] = "";

int ecmc_fail(int status, const char *msg) {
    snprintf(g_last_error, sizeof(g_last_error), "%s", msg ? msg : "");
    return status;
}

extern "C" const char *ecmc_last_error(void) { return g_last_error; }

namespace ecmc {

static int trunc_int(double x) { return (int)x; }

int CellGrid::position_to_cell(const double *p) const {
    long idx = 0;
    for (int d = 0; d < dim; d++) idx += (long)trunc_int(p[d] / side[d]) * cum[d];
    if (idx < 0 || idx >= count) return -1;
    return (int)idx;
}

bool build_cell_grid(int dim, double L, const int32_t *cells_per_side, int layers, CellGrid &g, std::string &err) {
    g.dim = dim;
    for (int d = 0; d < dim; d++) {
        int c = cells_per_side[d] > 0 ? cells_per_side[d] : cells_per_side[0];
        if (c <= 0) { err = "cells_per_side must be positive"; return false; }
        g.cps[d] = c;
        g.side[d] = L / c;
    }
    g.cum[0] = 1;
    for (int d = 1; d < dim; d++) g.cum[d] = g.cum[d - 1] * g.cps[d - 1];
    g.count = (int)(g.cum[dim - 1] * g.cps[dim - 1]);
    g.cmin.assign((size_t)g.count * dim, 0.0);
    g.cmax.assign((size_t)g.count * dim, 0.0);
    for (int cell = 0; cell < g.count; cell++) {
        for (int d = 0; d < dim; d++) {
            int id = (int)((cell / g.cum[d]) % g.cps[d]);
            double lo = id * g.side[d];
            double hi = (id + 1) * g.side[d];
            if (lo > 0.0) {
                while (trunc_int(lo / g.side[d]) == id) lo = nextafter(lo, -INFINITY);
                while (trunc_int(lo / g.side[d]) < id) lo = nextafter(lo, INFINITY);
            }
            while (trunc_int(hi / g.side[d]) == id) hi = nextafter(hi, INFINITY);
            while (trunc_int(hi / g.side[d]) > id) hi = nextafter(hi, -INFINITY);
            g.cmin[(size_t)cell * dim + d] = lo;
            g.cmax[(size_t)cell * dim + d] = hi;
        }
    }
    int span = 2 * layers + 1, per = 1;
    for (int d = 0; d < dim; d++) per *= span;
    g.nb_stride = per;
    g.nearby.assign((size_t)g.count * per, -1);
    g.nearby_n.assign(g.count, 0);
    for (int cell = 0; cell < g.count; cell++) {
        int id[3] = {0, 0, 0};
        for (int d = 0; d < dim; d++) id[d] = (int)((cell / g.cum[d]) % g.cps[d]);
        int32_t *out = g.nearby.data() + (size_t)cell * per;
        int n = 0;
        for (int t = 0; t < per; t++) {
            int rem = t, off[3] = {0, 0, 0};
            for (int d = dim - 1; d >= 0; d--) {
                off[d] = rem % span - layers;
                rem /= span;
            }
            long idx = 0;
            for (int d = 0; d < dim; d++) {
                int v = id[d] + off[d];
                v = ((v % g.cps[d]) + g.cps[d]) % g.cps[d];
                idx += v * g.cum[d];
            }
            bool dup = false;
            for (int q = 0; q < n; q++)
                if (out[q] == idx) dup = true;
            if (!dup) out[n++] = (int32_t)idx;
        }
        g.nearby_n[cell] = n;
    }
    return true;
}

void build_ewald_table(int fc, double alpha, double L, std::vector<double> &f) {
    const int n = fc + 1;
    f.assign((size_t)n * n * n, 0.0);
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 1; i < n; i++) {
                double coefficient = (j == 0 && k == 0) ? 1.0 : ((k == 0 || j == 0) ? 2.0 : 4.0);
                double norm_sq = i * i + j * j + k * k;
                f[((size_t)i * n + j) * n + k] =
                    4.0 * i * coefficient / (norm_sq * L * L) * exp(-M_PI * M_PI * norm_sq / (alpha * alpha));
            }
}

}  // namespace ecmc
