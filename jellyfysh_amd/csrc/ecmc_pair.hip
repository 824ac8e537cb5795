// ecmc_pair.hip -- batched per-pair evaluations: one thread per (separation, potential change)
// row, reading SoA-free row-major inputs straight from device memory.  These are the
// device counterparts of the reference's per-call potential methods (displacement and
// derivative), used by the plugin classes and by the parity tests.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>
#include <algorithm>
#include <vector>

#include "../../include/jellyfysh_amd.h"
#include "ecmc_common.h"
#include "ecmc_math.h"
#include "ecmc_potentials.h"
#include "ecmc_fp32_bounds.h"

namespace ecmc {

constexpr int PT = 256;

struct PairArgs {
    int32_t potential, dim;
    int64_t n;
    double par[4];
    double L;
    const int32_t *dir;
    const double *speed, *sep, *c1, *c2, *dU;
    double *out;
    Ewald ewald;  // merged-image Coulomb
};

template <int D>
__device__ void pair_row(const PairArgs &a, int64_t i, bool derivative) {
    double s[D];
#pragma unroll
    for (int d = 0; d < D; d++) s[d] = a.sep[i * D + d];
    const int dir = a.dir[i];
    const double speed = a.speed ? a.speed[i] : 1.0;
    const double c1 = a.c1 ? a.c1[i] : 1.0, c2 = a.c2 ? a.c2[i] : 1.0;
    const double dU = a.dU ? a.dU[i] : 0.0;
    double r = NAN;
    double p[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < 3; d++) p[d] = d < D ? s[(dir + d) % D] : 0.0;
    switch (a.potential) {
    case POT_HARD_SPHERE:
        r = derivative ? NAN : hs_displacement<D>(dir, speed, s, a.par[0]);
        break;
    case POT_HARD_DIPOLE:
        r = derivative ? NAN : hd_displacement<D>(dir, speed, s, a.par[0], a.par[1]);
        break;
    case POT_INVERSE_POWER:
        r = derivative ? ip_derivative<D>(dir, speed, s, c1, c2, a.par[0], a.par[1])
                       : ip_displacement<D>(dir, speed, s, c1, c2, a.par[0], a.par[1], dU);
        break;
    case POT_INVERSE_POWER_TRUNC:
        r = derivative ? ipt_derivative<D>(dir, speed, s, c1, c2, a.par[0], a.par[1], a.par[2])
                       : ipt_displacement<D>(dir, speed, s, c1, c2, a.par[0], a.par[1], a.par[2], dU);
        break;
    case POT_LENNARD_JONES: {
        MexicanHat h = mh_lennard_jones(a.par[0], a.par[1]);
        r = derivative ? mh_derivative<D>(h, dir, speed, s) : mh_displacement<D>(h, dir, speed, s, dU);
        break;
    }
    case POT_DISPLACED_EVEN_POWER: {
        MexicanHat h = mh_displaced_even_power(a.par[0], a.par[1], a.par[2]);
        r = derivative ? mh_derivative<D>(h, dir, speed, s) : mh_displacement<D>(h, dir, speed, s, dU);
        break;
    }
    case POT_INVERSE_POWER_COULOMB_BOUND: {
        const double pp = a.par[0] * c1 * c2;
        r = derivative ? ipcb_derivative(pp, p[0], p[1], p[2]) * speed
                       : ipcb_displacement(pp, p[0], p[1], p[2], dU, a.L) / speed;
        break;
    }
    case POT_MERGED_IMAGE_COULOMB:
        r = derivative ? a.par[3] * c1 * c2 * ewald_derivative(a.ewald, p[0], p[1], p[2]) * speed : NAN;
        break;
    default:
        break;
    }
    a.out[i] = r;
}

template <int D>
__global__ __launch_bounds__(PT) void pair_kernel(PairArgs a, int derivative) {
    int64_t i = (int64_t)blockIdx.x * PT + threadIdx.x;
    if (i >= a.n) return;
    pair_row<D>(a, i, derivative != 0);
}

// FP32 brackets of the time displacement of inverse-power rows and of inverse-power Coulomb
// bound rows (the chain kernel's candidate filter), the potential change drawn from the leading
// word of the random pair.
template <int D>
__global__ __launch_bounds__(PT) void pair_bounds_kernel(int64_t n, int trunc, int ipcb, double power, double prefactor,
                                                          double cutoff, double L, const int32_t *dir, const double *speed,
                                                          const double *sep, const double *c1, const double *c2,
                                                          const uint32_t *word, double beta, float *lb, float *ub) {
    const int64_t i = (int64_t)blockIdx.x * PT + threadIdx.x;
    if (i >= n) return;
    double s[D];
#pragma unroll
    for (int d = 0; d < D; d++) s[d] = sep[i * D + d];
    const double pp = prefactor * ((c1 ? c1[i] : 1.0) * (c2 ? c2[i] : 1.0));
    const float K = (float)fabs(pp);
    const float ib = (float)(1.0 / beta), is = (float)(1.0 / (speed ? speed[i] : 1.0));
    const float rc2 = trunc ? (float)(cutoff * cutoff) : 0.f;
    float l = 0.f, u = __int_as_float(0x7f800000);
    if (K >= 0x1p-100f && K <= 0x1p100f) {
        const float dU_lo = dU_lower(word[i], ib * (1.f - 0x1p-22f)), dU_hi = dU_upper(word[i], ib * (1.f + 0x1p-22f));
        if (ipcb) {
            float lo2 = 0.f, hi2 = 0.f;
#pragma unroll
            for (int d = 0; d < D; d++) {
                if (d == dir[i]) continue;
                const float m = fabsf((float)s[d]);
                const float lo = fmaxf(dn(m) - 0x1p-110f, 0.f), hi = up(m) + 0x1p-110f;
                lo2 += lo * lo;
                hi2 += hi * hi;
            }
            ipcb_time_bounds(dn(K), up(K), dn(lo2), up(hi2), (float)L, dU_lo, dU_hi, is * (1.f - 0x1p-22f),
                             is * (1.f + 0x1p-22f), l, u);
        } else {
            ip_time_bounds<D>(s, dir[i], K, pp > 0.0, (int)power, 2.f / (float)(int)power, trunc != 0, dn(rc2), up(rc2),
                              dU_lo, dU_hi, is * (1.f - 0x1p-22f), is * (1.f + 0x1p-22f), l, u);
        }
    }
    lb[i] = l;
    ub[i] = u;
}

__global__ __launch_bounds__(PT) void log_kernel(int64_t n, const double *x, double *out) {
    int64_t i = (int64_t)blockIdx.x * PT + threadIdx.x;
    if (i < n) out[i] = cr_log(x[i]);
}
__global__ __launch_bounds__(PT) void pow_kernel(int64_t n, const double *x, const double *y, double *out) {
    int64_t i = (int64_t)blockIdx.x * PT + threadIdx.x;
    if (i < n) out[i] = cr_pow(x[i], y[i]);
}

static int launch_pair(int32_t potential, const double *par, int32_t dim, int64_t n, const int32_t *dir,
                       const double *speed, const double *sep, const double *c1, const double *c2, const double *dU,
                       double L, double *out, void *stream, int derivative) {
    if (n < 0 || (dim != 2 && dim != 3) || !par) return ecmc_fail(ECMC_ERR_ARGUMENT, "bad argument");
    if (n == 0) return ECMC_OK;
    if (!dir || !sep || !out) return ecmc_fail(ECMC_ERR_ARGUMENT, "null device pointer");
    if (potential < 0 || potential > POT_HARD_DIPOLE || potential == POT_CELL_BOUNDING)
        return ecmc_fail(ECMC_ERR_CONFIGURATION, "unknown potential");
    if ((potential == POT_MERGED_IMAGE_COULOMB || potential == POT_INVERSE_POWER_COULOMB_BOUND) && dim != 3)
        return ecmc_fail(ECMC_ERR_CONFIGURATION, "the Coulomb potentials need three dimensions");
    if (potential == POT_MERGED_IMAGE_COULOMB && !derivative)
        return ecmc_fail(ECMC_ERR_CONFIGURATION, "the merged-image Coulomb potential has no displacement");
    if ((potential == POT_HARD_SPHERE || potential == POT_HARD_DIPOLE) && derivative)
        return ecmc_fail(ECMC_ERR_CONFIGURATION, "the hard-sphere and hard-dipole potentials have no derivative");
    PairArgs a;
    memset(&a, 0, sizeof(a));
    a.potential = potential;
    a.dim = dim;
    a.n = n;
    for (int k = 0; k < 4; k++) a.par[k] = par[k];
    a.L = L;
    a.dir = dir; a.speed = speed; a.sep = sep; a.c1 = c1; a.c2 = c2; a.dU = dU; a.out = out;
    hipStream_t st = (hipStream_t)stream;
    double *d_table = nullptr;
    if (potential == POT_MERGED_IMAGE_COULOMB) {
        std::vector<double> table;
        int fc = (int)par[1], pc = (int)par[2];
        if (fc < 0 || pc < 0 || !(par[0] > 0.0)) return ecmc_fail(ECMC_ERR_CONFIGURATION, "bad Ewald parameters");
        build_ewald_table(fc, par[0], L, table);
        if (hipMalloc((void **)&d_table, sizeof(double) * table.size()) != hipSuccess)
            return ecmc_fail(ECMC_ERR_HIP, "hipMalloc failed");
        hipMemcpyAsync(d_table, table.data(), sizeof(double) * table.size(), hipMemcpyHostToDevice, st);
        a.ewald.f = d_table;
        a.ewald.fc = fc; a.ewald.pc = pc; a.ewald.L = L;
        a.ewald.aol = par[0] / L; a.ewald.aol2 = par[0] * par[0] / (L * L);
        a.ewald.two_aol_rpi = 2.0 * par[0] / (L * sqrt(M_PI));
        a.ewald.tpol = 2.0 * M_PI / L;
    }
    dim3 grid((unsigned)((n + PT - 1) / PT));
    if (dim == 2) hipLaunchKernelGGL(pair_kernel<2>, grid, dim3(PT), 0, st, a, derivative);
    else hipLaunchKernelGGL(pair_kernel<3>, grid, dim3(PT), 0, st, a, derivative);
    hipError_t e = hipGetLastError();
    if (d_table) {
        hipStreamSynchronize(st);
        hipFree(d_table);
    }
    if (e != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "pair kernel launch failed");
    return ECMC_OK;
}

extern "C" int ecmc_pair_time_bounds(int32_t potential, const double *par, int32_t dim, int64_t n,
                                     const int32_t *d_dir, const double *d_speed, const double *d_sep,
                                     const double *d_c1, const double *d_c2, const uint32_t *d_word, double beta,
                                     float *d_lb, float *d_ub, void *stream) {
    if (n < 0 || (dim != 2 && dim != 3) || !par || !(beta > 0.0)) return ecmc_fail(ECMC_ERR_ARGUMENT, "bad argument");
    const int ipcb = potential == POT_INVERSE_POWER_COULOMB_BOUND;
    if (potential != POT_INVERSE_POWER && potential != POT_INVERSE_POWER_TRUNC && !ipcb)
        return ecmc_fail(ECMC_ERR_CONFIGURATION, "time brackets exist for the inverse-power potentials and the Coulomb bound only");
    if (ipcb && (dim != 3 || !(par[1] > 0.0)))
        return ecmc_fail(ECMC_ERR_CONFIGURATION, "the Coulomb-bound brackets need three dimensions and par[1] = system length > 0");
    if (!ipcb && !(par[0] >= 1.0 && par[0] <= 32.0 && par[0] == floor(par[0])))
        return ecmc_fail(ECMC_ERR_CONFIGURATION, "time brackets need an integer power from 1 to 32");
    if (n == 0) return ECMC_OK;
    if (!d_dir || !d_sep || !d_word || !d_lb || !d_ub) return ecmc_fail(ECMC_ERR_ARGUMENT, "null device pointer");
    const int trunc = potential == POT_INVERSE_POWER_TRUNC;
    const double cutoff = trunc ? par[2] : 0.0;
    // the inverse-power rows: par = (power, prefactor[, cutoff]); the Coulomb bound: (prefactor, L)
    const double power = ipcb ? 1.0 : par[0], prefactor = ipcb ? par[0] : par[1], L = ipcb ? par[1] : 0.0;
    dim3 grid((unsigned)((n + PT - 1) / PT));
    hipStream_t st = (hipStream_t)stream;
    if (dim == 2)
        hipLaunchKernelGGL(pair_bounds_kernel<2>, grid, dim3(PT), 0, st, n, trunc, ipcb, power, prefactor, cutoff, L, d_dir,
                           d_speed, d_sep, d_c1, d_c2, d_word, beta, d_lb, d_ub);
    else
        hipLaunchKernelGGL(pair_bounds_kernel<3>, grid, dim3(PT), 0, st, n, trunc, ipcb, power, prefactor, cutoff, L, d_dir,
                           d_speed, d_sep, d_c1, d_c2, d_word, beta, d_lb, d_ub);
    if (hipGetLastError() != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "bounds kernel launch failed");
    return ECMC_OK;
}

__global__ __launch_bounds__(PT) void ewald_kernel(Ewald ew, int64_t n, const double *sep, double *out) {
    int64_t i = (int64_t)blockIdx.x * PT + threadIdx.x;
    if (i < n) out[i] = ewald_derivative(ew, sep[3 * i], sep[3 * i + 1], sep[3 * i + 2]);
}

}  // namespace ecmc

using namespace ecmc;

// ---- merged-image Coulomb handle (merged_image_coulomb_potential.h:37-44) --------------------
struct ecmc_ewald {
    int fc = 0, pc = 0;
    double alpha = 0.0, L = 0.0;
    double *d_table = nullptr;
    size_t table_n = 0;
    int device = 0;
};

static Ewald ewald_of(const ecmc_ewald *e) {
    Ewald w;
    w.f = e->d_table;
    w.fc = e->fc; w.pc = e->pc; w.L = e->L;
    w.aol = e->alpha / e->L; w.aol2 = e->alpha * e->alpha / (e->L * e->L);
    w.two_aol_rpi = 2.0 * e->alpha / (e->L * sqrt(M_PI));
    w.tpol = 2.0 * M_PI / e->L;
    return w;
}

extern "C" int ecmc_ewald_create(int32_t fourier_cutoff, int32_t position_cutoff, double alpha, double system_length,
                                 ecmc_ewald **out) {
    if (!out) return ecmc_fail(ECMC_ERR_ARGUMENT, "null argument");
    *out = nullptr;
    if (fourier_cutoff < 0 || position_cutoff < 0 || !(alpha > 0.0) || !(system_length > 0.0))
        return ecmc_fail(ECMC_ERR_CONFIGURATION, "bad Ewald parameters");
    ecmc_ewald *e = new ecmc_ewald();
    e->fc = fourier_cutoff; e->pc = position_cutoff; e->alpha = alpha; e->L = system_length;
    std::vector<double> table;
    build_ewald_table(fourier_cutoff, alpha, system_length, table);
    e->table_n = table.size();
    if (hipGetDevice(&e->device) != hipSuccess ||
        hipMalloc((void **)&e->d_table, sizeof(double) * table.size()) != hipSuccess ||
        hipMemcpy(e->d_table, table.data(), sizeof(double) * table.size(), hipMemcpyHostToDevice) != hipSuccess) {
        ecmc_ewald_destroy(e);
        return ecmc_fail(ECMC_ERR_HIP, "Ewald table upload failed");
    }
    *out = e;
    return ECMC_OK;
}

extern "C" int ecmc_ewald_copy(const ecmc_ewald *src, ecmc_ewald **out) {
    if (!src || !out) return ecmc_fail(ECMC_ERR_ARGUMENT, "null argument");
    *out = nullptr;
    ecmc_ewald *e = new ecmc_ewald(*src);
    e->d_table = nullptr;
    if (hipMalloc((void **)&e->d_table, sizeof(double) * src->table_n) != hipSuccess ||
        hipMemcpy(e->d_table, src->d_table, sizeof(double) * src->table_n, hipMemcpyDeviceToDevice) != hipSuccess) {
        ecmc_ewald_destroy(e);
        return ecmc_fail(ECMC_ERR_HIP, "Ewald table copy failed");
    }
    *out = e;
    return ECMC_OK;
}

extern "C" void ecmc_ewald_destroy(ecmc_ewald *e) {
    if (!e) return;
    hipFree(e->d_table);
    delete e;
}

extern "C" int ecmc_ewald_size(const ecmc_ewald *e, int64_t *bytes) {
    if (!e || !bytes) return ecmc_fail(ECMC_ERR_ARGUMENT, "null argument");
    *bytes = (int64_t)(sizeof(ecmc_ewald) + sizeof(double) * e->table_n);
    return ECMC_OK;
}

extern "C" int ecmc_ewald_derivative(const ecmc_ewald *e, int64_t n, const double *d_sep, double *d_out, void *stream) {
    if (!e || n < 0) return ecmc_fail(ECMC_ERR_ARGUMENT, "bad argument");
    if (n == 0) return ECMC_OK;
    if (!d_sep || !d_out) return ecmc_fail(ECMC_ERR_ARGUMENT, "null device pointer");
    hipLaunchKernelGGL(ewald_kernel, dim3((unsigned)((n + PT - 1) / PT)), dim3(PT), 0, (hipStream_t)stream, ewald_of(e),
                       n, d_sep, d_out);
    if (hipGetLastError() != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "Ewald kernel launch failed");
    return ECMC_OK;
}

extern "C" int ecmc_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" int ecmc_pair_displacement(int32_t potential, const double *par, int32_t dim, int64_t n,
                                      const int32_t *d_dir, const double *d_speed, const double *d_sep,
                                      const double *d_c1, const double *d_c2, const double *d_dU,
                                      double system_length, double *d_out, void *stream) {
    return launch_pair(potential, par, dim, n, d_dir, d_speed, d_sep, d_c1, d_c2, d_dU, system_length, d_out, stream, 0);
}

extern "C" int ecmc_pair_derivative(int32_t potential, const double *par, int32_t dim, int64_t n,
                                    const int32_t *d_dir, const double *d_speed, const double *d_sep,
                                    const double *d_c1, const double *d_c2, double system_length, double *d_out,
                                    void *stream) {
    return launch_pair(potential, par, dim, n, d_dir, d_speed, d_sep, d_c1, d_c2, nullptr, system_length, d_out,
                       stream, 1);
}

__global__ __launch_bounds__(PT) void copy_kernel(const uint4 *src, uint4 *dst, int64_t n16) {
    for (int64_t i = (int64_t)blockIdx.x * PT + threadIdx.x; i < n16; i += (int64_t)gridDim.x * PT) dst[i] = src[i];
}

extern "C" int ecmc_copy_bytes(const void *d_src, void *d_dst, int64_t bytes, void *stream) {
    if (bytes < 0 || (bytes % 16) != 0 || (bytes && (!d_src || !d_dst))) return ecmc_fail(ECMC_ERR_ARGUMENT, "bad argument");
    if (bytes == 0) return ECMC_OK;
    int64_t n16 = bytes / 16;
    int64_t blocks = std::min<int64_t>((n16 + PT - 1) / PT, 4096);
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(PT), 0, (hipStream_t)stream, (const uint4 *)d_src,
                       (uint4 *)d_dst, n16);
    return hipGetLastError() == hipSuccess ? ECMC_OK : ecmc_fail(ECMC_ERR_HIP, "copy kernel launch failed");
}

// Largest error of the FP32 log2f (absolute below 1, relative above: |err| / max(|log2 x|, 1))
// or exp2f (relative) that the batched kernel's FP32 bounds rely on, against FP64 log2 /
// exp2, over every float whose bit pattern lies in [lo, hi).  The result is a non-negative double, so its bit pattern orders like an unsigned
// integer and one atomicMax per wavefront combines the lanes.
__global__ __launch_bounds__(PT) void fp32_math_error_kernel(int fn, uint32_t lo, uint32_t hi,
                                                             unsigned long long *out) {
    double worst = 0.0;
    for (uint64_t b = (uint64_t)lo + (uint64_t)blockIdx.x * PT + threadIdx.x; b < hi;
         b += (uint64_t)gridDim.x * PT) {
        const float x = __uint_as_float((uint32_t)b);
        double err;
        if (fn == 0) {
            const double got = (double)log2f(x), ref = log2((double)x);
            err = fabs(got - ref) / fmax(fabs(ref), 1.0);
        } else {
            const double got = (double)exp2f(x), ref = exp2((double)x);
            err = fabs(got - ref) / ref;
        }
        if (!(err <= worst)) worst = err;      // NaN propagates as the worst case
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(worst, off, 64);
        if (!(o <= worst)) worst = o;
    }
    if ((threadIdx.x & 63) == 0) {
        unsigned long long bits = __double_as_longlong(worst);
        if (worst != worst) bits = 0x7ff8000000000000ull;
        atomicMax(out, bits);
    }
}

extern "C" int ecmc_fp32_math_error(int32_t fn, uint32_t bits_lo, uint32_t bits_hi, unsigned long long *d_out,
                                    void *stream) {
    if ((fn != 0 && fn != 1) || bits_hi < bits_lo || !d_out) return ecmc_fail(ECMC_ERR_ARGUMENT, "bad argument");
    if (bits_hi == bits_lo) return ECMC_OK;
    hipLaunchKernelGGL(fp32_math_error_kernel, dim3(4096), dim3(PT), 0, (hipStream_t)stream, fn, bits_lo, bits_hi,
                       d_out);
    return hipGetLastError() == hipSuccess ? ECMC_OK : ecmc_fail(ECMC_ERR_HIP, "math check launch failed");
}

// Norms of the shortest separations of every pair i < j, in the order SeparationOutputHandler
// writes them for point masses (output_handler/separation_output_handler.py:81-95): separation
// = target - reference with each entry ((s + L/2) % L) - L/2 (hypercubic_setting.py:120-172),
// norm = (0 + sum of squares) ** 0.5 (vectors.py:29-43), ** 0.5 being the correctly rounded
// square root.  Row i is blockIdx.y; pair (i, j) lands at i n - i (i + 1) / 2 + j - i - 1.
template <int D>
__global__ void separation_norm_kernel(int64_t n, const double *pos, double L, double *out) {
    const int64_t i = blockIdx.y;
    const int64_t j = i + 1 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double half = L / 2.0;
    double acc = 0.0;
#pragma unroll
    for (int d = 0; d < D; d++) {
        const double sd = py_mod((pos[D * j + d] - pos[D * i + d]) + half, L) - half;
        acc = acc + sd * sd;
    }
    out[i * n - i * (i + 1) / 2 + (j - i - 1)] = cr_pow(acc, 0.5);
}

extern "C" int ecmc_separation_norms(int32_t dim, int64_t n, const double *d_positions, double system_length,
                                     double *d_out, void *stream) {
    if (!(dim == 2 || dim == 3) || n < 0 || n > 65535 || !(system_length > 0.0))
        return ecmc_fail(ECMC_ERR_ARGUMENT, "bad argument");
    if (n < 2) return ECMC_OK;
    if (!d_positions || !d_out) return ecmc_fail(ECMC_ERR_ARGUMENT, "null argument");
    const dim3 grid((unsigned)((n - 1 + PT - 1) / PT), (unsigned)(n - 1));
    if (dim == 2)
        hipLaunchKernelGGL(separation_norm_kernel<2>, grid, dim3(PT), 0, (hipStream_t)stream, n, d_positions,
                           system_length, d_out);
    else
        hipLaunchKernelGGL(separation_norm_kernel<3>, grid, dim3(PT), 0, (hipStream_t)stream, n, d_positions,
                           system_length, d_out);
    return hipGetLastError() == hipSuccess ? ECMC_OK : ecmc_fail(ECMC_ERR_HIP, "separation kernel launch failed");
}

extern "C" int ecmc_math_log(int64_t n, const double *d_x, double *d_out, void *stream) {
    if (n < 0) return ecmc_fail(ECMC_ERR_ARGUMENT, "bad argument");
    if (n == 0) return ECMC_OK;
    hipLaunchKernelGGL(log_kernel, dim3((unsigned)((n + PT - 1) / PT)), dim3(PT), 0, (hipStream_t)stream, n, d_x, d_out);
    return hipGetLastError() == hipSuccess ? ECMC_OK : ecmc_fail(ECMC_ERR_HIP, "log kernel launch failed");
}

extern "C" int ecmc_math_pow(int64_t n, const double *d_x, const double *d_y, double *d_out, void *stream) {
    if (n < 0) return ecmc_fail(ECMC_ERR_ARGUMENT, "bad argument");
    if (n == 0) return ECMC_OK;
    hipLaunchKernelGGL(pow_kernel, dim3((unsigned)((n + PT - 1) / PT)), dim3(PT), 0, (hipStream_t)stream, n, d_x, d_y,
                       d_out);
    return hipGetLastError() == hipSuccess ? ECMC_OK : ecmc_fail(ECMC_ERR_HIP, "pow kernel launch failed");
}
