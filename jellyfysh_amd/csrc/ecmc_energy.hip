// ecmc_energy.hip -- total pair energy of one configuration and its hand-written backward
// kernel: gradients with respect to the positions, to the potential's parameters and to
// the charges.
//
// U = 1/2 sum_i sum_{j != i, (i, j) in scope} U(s_ij), s_ij = minimum-image separation
// (hypercubic_setting.py:101-126).  The position gradient of a pair is the potential's own
// directional derivative -- the same device functions ECMC inverts
// (inverse_power_potential.py:71-94, lennard_jones_potential.py:63-81,
// displaced_even_power_potential.py:73-93, merged_image_coulomb_potential.c:205-274) -- so
// dU/dx_{i,d} = sum_j derivative(e_d, s_ij, c_i, c_j).
//
// One workgroup per row i: its 256 threads stride over j in a fixed order and reduce in a
// fixed tree, and a last kernel sums the rows in a fixed order, so results do not depend on
// scheduling.  All scratch lives in a caller-provided workspace; nothing is allocated here.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>
#include <vector>

#include "../../include/jellyfysh_amd.h"
#include "ecmc_common.h"
#include "ecmc_math.h"
#include "ecmc_potentials.h"

namespace ecmc {

constexpr int ET = 256;
constexpr int EROW = 9;   // energy, dU/dx_i (3), dU/dpar (4), dU/dc_i

struct EnergyArgs {
    int32_t pot, scope, use_charge, n;
    double par[4];
    double L, half, background;
    const double *pos;
    const double *charge;
    const int32_t *mol;
    Ewald ew;            // derivative table (merged_image_coulomb_potential.c:94-109)
    const double *g;     // Fourier table of the pair function itself
    double *rows;        // n x EROW
};

// Merged-image Coulomb pair function psi(s) = sum over the real-space images of the
// derivative's sphere of erfc(a |s + nL|) / |s + nL| + (1 / (pi L)) sum_{0 < |n| <= fc}
// exp(-pi^2 n^2 / alpha^2) / n^2 cos(2 pi n.s / L) - pi / (alpha^2 L), a = alpha / L.
// Its gradient is the reference's derivative; the last term (the neutralising background)
// makes psi independent of alpha up to truncation.
__device__ double ewald_psi(const EnergyArgs &A, double sx, double sy, double sz) {
    const Ewald &e = A.ew;
    double psi = 0.0;
    const int pc = e.pc, pc2 = pc * pc, fc = e.fc, fc2 = fc * fc, nn = fc + 1;
    const double L = e.L;
    for (int k = -pc; k < pc + 1; k++) {
        double vz2 = (sz + k * L) * (sz + k * L);
        int cy = (int)sqrt((double)(pc2 - k * k));
        for (int j = -cy; j < cy + 1; j++) {
            double vy2 = (sy + j * L) * (sy + j * L);
            int cx = (int)sqrt((double)(pc2 - j * j - k * k));
            for (int i = -cx; i < cx + 1; i++) {
                double vx = sx + i * L;
                double vn = sqrt(vx * vx + vy2 + vz2);
                psi += erfc(e.aol * vn) / vn;
            }
        }
    }
    for (int i = 0; i <= fc; i++) {
        double ci = cos(e.tpol * i * sx);
        int cy = (int)sqrt((double)(fc2 - i * i));
        for (int j = 0; j <= cy; j++) {
            double cij = ci * cos(e.tpol * j * sy);
            int cx = (int)sqrt((double)(fc2 - i * i - j * j));
            for (int k = 0; k <= cx; k++) {
                if (i == 0 && j == 0 && k == 0) continue;
                psi += A.g[(i * nn + j) * nn + k] * cij * cos(e.tpol * k * sz);
            }
        }
    }
    return psi + A.background;
}

// Pair value u (charges included), its derivative along each axis for the row unit, the
// parameter gradient and du/dc_i.
template <bool BW>
__device__ void pair_terms(const EnergyArgs &A, const double *s, double ci, double cj, double &u, double *dx,
                           double *dp, double &dc) {
    const double cp = ci * cj;
    u = 0.0;
    dc = 0.0;
    if (BW) {
        for (int q = 0; q < 3; q++) dx[q] = 0.0;
        for (int q = 0; q < 4; q++) dp[q] = 0.0;
    }
    const double r2 = norm_sq<3>(s);
    switch (A.pot) {
    case ECMC_POT_INVERSE_POWER:
    case ECMC_POT_INVERSE_POWER_TRUNC: {
        const double p = A.par[0], k = A.par[1];
        const bool trunc = A.pot == ECMC_POT_INVERSE_POWER_TRUNC;
        const double rc = A.par[2];
        if (trunc && !(r2 < rc * rc)) return;
        const double rp = 1.0 / cr_pow(r2, p / 2.0);                 // r^-p
        const double rcp = trunc ? 1.0 / cr_pow(rc * rc, p / 2.0) : 0.0;
        u = cp * k * (rp - rcp);
        if (BW) {
            for (int d = 0; d < 3; d++) dx[d] = ip_derivative<3>(d, 1.0, s, ci, cj, p, k);
            const double lr = 0.5 * cr_log(r2);
            dp[0] = -cp * k * (rp * lr - (trunc ? rcp * cr_log(rc) : 0.0));
            dp[1] = cp * (rp - rcp);
            if (trunc) dp[2] = cp * k * p * rcp / rc;
            dc = cj * k * (rp - rcp);
        }
        return;
    }
    case ECMC_POT_LENNARD_JONES: {
        const double k = A.par[0], sig = A.par[1];
        const MexicanHat h = mh_lennard_jones(k, sig);
        u = mh_potential<3>(h, s);
        if (BW) {
            for (int d = 0; d < 3; d++) dx[d] = mh_derivative<3>(h, d, 1.0, s);
            const double x6 = cr_pow(sig * sig / r2, 3.0), x12 = x6 * x6;
            dp[0] = x12 - x6;
            dp[1] = k * (12.0 * x12 - 6.0 * x6) / sig;
        }
        return;
    }
    case ECMC_POT_DISPLACED_EVEN_POWER: {
        const double r0 = A.par[0], p = A.par[1], k = A.par[2];
        const MexicanHat h = mh_displaced_even_power(r0, p, k);
        u = mh_potential<3>(h, s);
        if (BW) {
            for (int d = 0; d < 3; d++) dx[d] = mh_derivative<3>(h, d, 1.0, s);
            const double dr = sqrt(r2) - r0;
            const double dpw = cr_pow(dr, p);
            dp[0] = -k * p * cr_pow(dr, p - 1.0);
            dp[1] = dr != 0.0 ? k * dpw * cr_log(fabs(dr)) : 0.0;
            dp[2] = dpw;
        }
        return;
    }
    case ECMC_POT_MERGED_IMAGE_COULOMB: {
        const double pre = A.par[3];
        const double psi = ewald_psi(A, s[0], s[1], s[2]);
        u = pre * cp * psi;
        if (BW) {
            for (int d = 0; d < 3; d++)
                dx[d] = pre * cp * ewald_derivative(A.ew, s[d], s[(d + 1) % 3], s[(d + 2) % 3]);
            dp[3] = cp * psi;
            dc = pre * cj * psi;
        }
        return;
    }
    default:
        return;
    }
}

template <bool BW>
__global__ __launch_bounds__(ET) void energy_rows(EnergyArgs A) {
    __shared__ double red[EROW][ET];
    const int i = blockIdx.x, t = threadIdx.x;
    double acc[EROW];
#pragma unroll
    for (int q = 0; q < EROW; q++) acc[q] = 0.0;
    const double ai[3] = {A.pos[3 * (size_t)i], A.pos[3 * (size_t)i + 1], A.pos[3 * (size_t)i + 2]};
    const double ci = A.charge ? A.charge[i] : 1.0;
    const int mi = A.mol ? A.mol[i] : 0;
    for (int j = t; j < A.n; j += ET) {
        if (j == i) continue;
        if (A.scope == ECMC_SCOPE_INTER && A.mol[j] == mi) continue;
        if (A.scope == ECMC_SCOPE_INTRA && A.mol[j] != mi) continue;
        double s[3];
#pragma unroll
        for (int d = 0; d < 3; d++) s[d] = py_mod((A.pos[3 * (size_t)j + d] - ai[d]) + A.half, A.L) - A.half;
        const double cj = A.use_charge ? A.charge[j] : 1.0;
        double u, dx[3], dp[4], dc;
        pair_terms<BW>(A, s, A.use_charge ? ci : 1.0, cj, u, dx, dp, dc);
        acc[0] += u;
        if (BW) {
#pragma unroll
            for (int q = 0; q < 3; q++) acc[1 + q] += dx[q];
#pragma unroll
            for (int q = 0; q < 4; q++) acc[4 + q] += dp[q];
            acc[8] += dc;
        }
    }
#pragma unroll
    for (int q = 0; q < EROW; q++) red[q][t] = acc[q];
    __syncthreads();
    for (int w = ET / 2; w > 0; w >>= 1) {
        if (t < w)
#pragma unroll
            for (int q = 0; q < EROW; q++) red[q][t] += red[q][t + w];
        __syncthreads();
    }
    if (t < EROW) A.rows[(size_t)i * EROW + t] = red[t][0];
}

// Fixed-order sums over the rows: energy (and parameter gradient) = half the row sums.
__global__ __launch_bounds__(ET) void energy_finish(const double *rows, int n, int backward, double *energy,
                                                     double *grad_pos, double *grad_par, double *grad_charge) {
    __shared__ double red[5][ET];
    const int t = threadIdx.x;
    double acc[5] = {0, 0, 0, 0, 0};
    for (int i = t; i < n; i += ET) {
        const double *r = rows + (size_t)i * EROW;
        acc[0] += r[0];
        if (backward) {
            for (int q = 0; q < 4; q++) acc[1 + q] += r[4 + q];
            for (int d = 0; d < 3; d++) grad_pos[3 * (size_t)i + d] = r[1 + d];
            if (grad_charge) grad_charge[i] = r[8];
        }
    }
    for (int q = 0; q < 5; q++) red[q][t] = acc[q];
    __syncthreads();
    for (int w = ET / 2; w > 0; w >>= 1) {
        if (t < w)
            for (int q = 0; q < 5; q++) red[q][t] += red[q][t + w];
        __syncthreads();
    }
    if (t == 0) {
        if (energy) energy[0] = 0.5 * red[0][0];
        if (backward)
            for (int q = 0; q < 4; q++) grad_par[q] = 0.5 * red[1 + q][0];
    }
}

}  // namespace ecmc

using namespace ecmc;

static int64_t fourier_entries(const ecmc_energy_term *t) {
    if (t->potential != ECMC_POT_MERGED_IMAGE_COULOMB) return 0;
    const int64_t nn = (int64_t)t->par[1] + 1;
    return nn * nn * nn;
}

extern "C" int ecmc_energy_workspace_bytes(const ecmc_energy_term *term, int32_t n, int64_t *bytes) {
    if (!term || !bytes || n < 0) return ecmc_fail(ECMC_ERR_ARGUMENT, "null argument or negative size");
    *bytes = (int64_t)n * EROW * 8 + 2 * fourier_entries(term) * 8;
    return ECMC_OK;
}

static int energy_launch(const ecmc_energy_term *term, int32_t n, const double *d_pos, const double *d_charge,
                         const int32_t *d_mol, double L, bool backward, double *d_energy, double *d_grad_pos,
                         double *d_grad_par, double *d_grad_charge, void *d_work, void *stream) {
    if (!term || !d_pos || !d_work) return ecmc_fail(ECMC_ERR_ARGUMENT, "null argument");
    if (n < 2) return ecmc_fail(ECMC_ERR_ARGUMENT, "at least two units");
    if (!(L > 0.0)) return ecmc_fail(ECMC_ERR_ARGUMENT, "system length must be positive");
    if (term->scope < ECMC_SCOPE_ALL || term->scope > ECMC_SCOPE_INTRA)
        return ecmc_fail(ECMC_ERR_ARGUMENT, "unknown pair scope");
    if (term->scope != ECMC_SCOPE_ALL && !d_mol) return ecmc_fail(ECMC_ERR_ARGUMENT, "scope needs molecule indices");
    if (term->use_charge && !d_charge) return ecmc_fail(ECMC_ERR_ARGUMENT, "charged term needs charges");
    if (backward && (!d_grad_pos || !d_grad_par)) return ecmc_fail(ECMC_ERR_ARGUMENT, "null gradient output");
    if (d_grad_charge && !term->use_charge)
        return ecmc_fail(ECMC_ERR_ARGUMENT, "charge gradient of an uncharged term");
    EnergyArgs A;
    memset(&A, 0, sizeof(A));
    A.pot = term->potential;
    A.scope = term->scope;
    A.use_charge = term->use_charge;
    A.n = n;
    for (int q = 0; q < 4; q++) A.par[q] = term->par[q];
    A.L = L;
    A.half = 0.5 * L;
    A.pos = d_pos;
    A.charge = term->use_charge ? d_charge : nullptr;
    A.mol = d_mol;
    A.rows = (double *)d_work;
    hipStream_t st = (hipStream_t)stream;
    switch (A.pot) {
    case ECMC_POT_INVERSE_POWER:
        if (!(A.par[0] > 0.0)) return ecmc_fail(ECMC_ERR_CONFIGURATION, "power must be positive");
        break;
    case ECMC_POT_INVERSE_POWER_TRUNC:
        if (!(A.par[0] > 0.0)) return ecmc_fail(ECMC_ERR_CONFIGURATION, "power must be positive");
        if (!(A.par[2] > 0.0) || A.par[2] > 0.5 * L)
            return ecmc_fail(ECMC_ERR_CONFIGURATION, "cutoff must lie in (0, L/2] (minimum image)");
        break;
    case ECMC_POT_LENNARD_JONES:
        break;
    case ECMC_POT_DISPLACED_EVEN_POWER: {
        double p = A.par[1];
        if (!(p > 0.0) || p != (double)(int)p || ((int)p) % 2)
            return ecmc_fail(ECMC_ERR_CONFIGURATION, "the power must be a positive even integer");
        break;
    }
    case ECMC_POT_MERGED_IMAGE_COULOMB: {
        const int fc = (int)A.par[1], pc = (int)A.par[2];
        const double alpha = A.par[0];
        if (fc < 1 || fc > 32 || pc < 0 || pc > 8 || !(alpha > 0.0))
            return ecmc_fail(ECMC_ERR_CONFIGURATION, "Ewald cutoffs out of range");
        std::vector<double> f, g;
        build_ewald_table(fc, alpha, L, f);
        const int nn = fc + 1;
        g.assign((size_t)nn * nn * nn, 0.0);
        for (int i = 0; i < nn; i++)
            for (int j = 0; j < nn; j++)
                for (int k = 0; k < nn; k++) {
                    int n2 = i * i + j * j + k * k;
                    if (n2 == 0 || n2 > fc * fc) continue;
                    double w = (i ? 2.0 : 1.0) * (j ? 2.0 : 1.0) * (k ? 2.0 : 1.0);
                    g[((size_t)i * nn + j) * nn + k] =
                        w * exp(-M_PI * M_PI * n2 / (alpha * alpha)) / (M_PI * L * n2);
                }
        double *df = A.rows + (size_t)n * EROW, *dg = df + f.size();
        if (hipMemcpyAsync(df, f.data(), f.size() * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(dg, g.data(), g.size() * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)   // f and g are host temporaries
            return ecmc_fail(ECMC_ERR_HIP, "Ewald table upload failed");
        A.ew.f = df;
        A.ew.fc = fc;
        A.ew.pc = pc;
        A.ew.L = L;
        A.ew.aol = alpha / L;
        A.ew.aol2 = A.ew.aol * A.ew.aol;
        A.ew.two_aol_rpi = 2.0 * A.ew.aol / sqrt(M_PI);
        A.ew.tpol = 2.0 * M_PI / L;
        A.g = dg;
        A.background = -M_PI / (alpha * alpha * L);
        break;
    }
    default:
        return ecmc_fail(ECMC_ERR_CONFIGURATION, "no energy for this potential");
    }
    if (backward) hipLaunchKernelGGL(energy_rows<true>, dim3(n), dim3(ET), 0, st, A);
    else hipLaunchKernelGGL(energy_rows<false>, dim3(n), dim3(ET), 0, st, A);
    hipLaunchKernelGGL(energy_finish, dim3(1), dim3(ET), 0, st, (const double *)A.rows, n, backward ? 1 : 0, d_energy,
                       d_grad_pos, d_grad_par, d_grad_charge);
    if (hipGetLastError() != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "energy kernel launch failed");
    return ECMC_OK;
}

extern "C" int ecmc_energy_forward(const ecmc_energy_term *term, int32_t n, const double *d_pos,
                                   const double *d_charge, const int32_t *d_molecule, double system_length,
                                   double *d_energy, void *d_work, void *stream) {
    if (!d_energy) return ecmc_fail(ECMC_ERR_ARGUMENT, "null energy output");
    return energy_launch(term, n, d_pos, d_charge, d_molecule, system_length, false, d_energy, nullptr, nullptr,
                         nullptr, d_work, stream);
}

extern "C" int ecmc_energy_backward(const ecmc_energy_term *term, int32_t n, const double *d_pos,
                                    const double *d_charge, const int32_t *d_molecule, double system_length,
                                    double *d_grad_pos, double *d_grad_par, double *d_grad_charge, void *d_work,
                                    void *stream) {
    return energy_launch(term, n, d_pos, d_charge, d_molecule, system_length, true, nullptr, d_grad_pos, d_grad_par,
                         d_grad_charge, d_work, stream);
}
