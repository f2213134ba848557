// Weight integrals of attractive-vector biasing on the GPU (FP64).
//
// ComputeWeightIntegral (base/xmcml/xmcml/mcml_integration.cpp:7-82) is a
// 2-point Gauss-Legendre rule on n x n cells of (theta, phi) in
// [0, pi] x [0, 2 pi] of
//     f_HG(x) * (0.1 + 0.45 (1 + px))^eps * sin(theta),
//     px = sin(theta) cos(phi),  x = dot px + sqrt(1 - dot^2) sin(theta) sin(phi).
// The reference evaluates every sine, cosine and power per point and per
// table entry (weight_integral.cpp:33-64: 262,145 entries x 4 n^2 points,
// hours of CPU).  Here the host evaluates the parts that do not depend on
// dot once, with libm, exactly as the reference writes them; the GPU runs one
// table entry per thread over the n^2 cells in the reference's order with the
// remaining IEEE operations (+, *, /, sqrt; no FMA contraction), so each entry
// equals the reference's value.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/xmcml.h"

namespace xmc {
int set_error(int code, const std::string& msg);
}

namespace {

constexpr double kPi = 3.14159265358979323846;

// Per cell row i: sin(theta) at the two nodes and the width; per column j:
// sin and cos(phi) at the two nodes and the width; per cell the four powers
// (0.1 + 0.45 (1 + px))^eps in the order (xp,yp), (xm,yp), (xp,ym), (xm,ym).
struct Rule {
    int n;
    const double* row;  // 3 per i: sin(xp), sin(xm), b - a
    const double* col;  // 5 per j: sin(yp), sin(ym), cos(yp), cos(ym), d - c
    const double* pw;   // 4 per cell (i * n + j)
};

__global__ void __launch_bounds__(256) weight_integral_kernel(Rule R, const double* __restrict__ dots, int m,
                                                             double g, double* __restrict__ out) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const double dot = dots[k];
    const double g2 = g * g;
    const double sqrt_dot = __builtin_sqrt(1 - dot * dot);
    const double one_m_g2 = 1 - g2, one_p_g2 = 1 + g2, two_g = 2 * g;
    double sum = 0.0;
    for (int i = 0; i < R.n; ++i) {
        const double sxp = R.row[3 * i], sxm = R.row[3 * i + 1], bma = R.row[3 * i + 2];
        for (int j = 0; j < R.n; ++j) {
            const double* c = R.col + 5 * j;
            const double syp = c[0], sym = c[1], cyp = c[2], cym = c[3], dmc = c[4];
            const double* p = R.pw + 4 * ((size_t)i * R.n + j);
            double px, fx, tmp, f;
            // func(xp, yp)
            px = sxp * cyp;
            fx = dot * px + sqrt_dot * sxp * syp;
            tmp = one_p_g2 - two_g * fx;
            f = one_m_g2 / (2 * tmp * __builtin_sqrt(tmp));
            const double f_xp_yp = f * p[0] * sxp;
            // func(xm, yp)
            px = sxm * cyp;
            fx = dot * px + sqrt_dot * sxm * syp;
            tmp = one_p_g2 - two_g * fx;
            f = one_m_g2 / (2 * tmp * __builtin_sqrt(tmp));
            const double f_xm_yp = f * p[1] * sxm;
            // func(xp, ym)
            px = sxp * cym;
            fx = dot * px + sqrt_dot * sxp * sym;
            tmp = one_p_g2 - two_g * fx;
            f = one_m_g2 / (2 * tmp * __builtin_sqrt(tmp));
            const double f_xp_ym = f * p[2] * sxp;
            // func(xm, ym)
            px = sxm * cym;
            fx = dot * px + sqrt_dot * sxm * sym;
            tmp = one_p_g2 - two_g * fx;
            f = one_m_g2 / (2 * tmp * __builtin_sqrt(tmp));
            const double f_xm_ym = f * p[3] * sxm;
            sum += 0.25 * bma * dmc * (f_xp_yp + f_xp_ym + f_xm_yp + f_xm_ym);
        }
    }
    out[k] = sum / (2.0 * kPi);
}

// The host half of the rule for n cells per axis and exponent eps
// (mcml_integration.cpp:20-76, the dot-independent operations).
void host_rule(int n, double eps, std::vector<double>& row, std::vector<double>& col, std::vector<double>& pw) {
    const double coeff = 2.0 * std::sqrt(3.0);
    const double xh = kPi / n, yh = 2 * kPi / n;
    row.resize(3 * (size_t)n);
    col.resize(5 * (size_t)n);
    pw.resize(4 * (size_t)n * n);
    std::vector<double> xp(n), xm(n), yp(n), ym(n);
    for (int i = 0; i < n; i++) {
        double a = xh * i, b = a + xh;
        xp[i] = 0.5 * (a + b) + (b - a) / coeff;
        xm[i] = 0.5 * (a + b) - (b - a) / coeff;
        row[3 * i] = std::sin(xp[i]);
        row[3 * i + 1] = std::sin(xm[i]);
        row[3 * i + 2] = b - a;
        double c = yh * i, d = c + yh;
        yp[i] = 0.5 * (c + d) + (d - c) / coeff;
        ym[i] = 0.5 * (c + d) - (d - c) / coeff;
        col[5 * i] = std::sin(yp[i]);
        col[5 * i + 1] = std::sin(ym[i]);
        col[5 * i + 2] = std::cos(yp[i]);
        col[5 * i + 3] = std::cos(ym[i]);
        col[5 * i + 4] = d - c;
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double* p = &pw[4 * ((size_t)i * n + j)];
            const double sxp = row[3 * i], sxm = row[3 * i + 1];
            const double cyp = col[5 * j + 2], cym = col[5 * j + 3];
            p[0] = std::pow(0.1 + 0.9 * 0.5 * (1 + sxp * cyp), eps);
            p[1] = std::pow(0.1 + 0.9 * 0.5 * (1 + sxm * cyp), eps);
            p[2] = std::pow(0.1 + 0.9 * 0.5 * (1 + sxp * cym), eps);
            p[3] = std::pow(0.1 + 0.9 * 0.5 * (1 + sxm * cym), eps);
        }
}

template <class T>
struct Dev {
    T* p = nullptr;
    ~Dev() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t n) { return hipMalloc(reinterpret_cast<void**>(&p), (n ? n : 1) * sizeof(T)); }
};

#define BIAS_TRY(expr)                                                                         \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            return xmc::set_error(XMC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

}  // namespace

extern "C" {

int xmc_weight_integrals(const double* dots, int m, double g, double eps, int n, double* out, int device) {
    if ((!dots && m > 0) || (!out && m > 0) || m < 0) return xmc::set_error(XMC_ERR_ARG, "dots / out");
    if (n < 1 || n > 65536) return xmc::set_error(XMC_ERR_LIMIT, "integral precision must be in [1, 65536]");
    if (m == 0) return XMC_OK;
    for (int k = 0; k < m; k++)
        if (!(dots[k] >= -1.0 && dots[k] <= 1.0)) return xmc::set_error(XMC_ERR_ARG, "dot outside [-1, 1]");
    std::vector<double> row, col, pw;
    host_rule(n, eps, row, col, pw);
    BIAS_TRY(hipSetDevice(device));
    Dev<double> drow, dcol, dpw, ddots, dout;
    BIAS_TRY(drow.alloc(row.size()));
    BIAS_TRY(dcol.alloc(col.size()));
    BIAS_TRY(dpw.alloc(pw.size()));
    BIAS_TRY(ddots.alloc(m));
    BIAS_TRY(dout.alloc(m));
    BIAS_TRY(hipMemcpy(drow.p, row.data(), row.size() * 8, hipMemcpyHostToDevice));
    BIAS_TRY(hipMemcpy(dcol.p, col.data(), col.size() * 8, hipMemcpyHostToDevice));
    BIAS_TRY(hipMemcpy(dpw.p, pw.data(), pw.size() * 8, hipMemcpyHostToDevice));
    BIAS_TRY(hipMemcpy(ddots.p, dots, (size_t)m * 8, hipMemcpyHostToDevice));
    Rule R{n, drow.p, dcol.p, dpw.p};
    const int block = 64;  // one wave per group: more groups than CUs for tables of >= 16k entries
    hipLaunchKernelGGL(weight_integral_kernel, dim3((m + block - 1) / block), dim3(block), 0, 0, R, ddots.p, m, g,
                       dout.p);
    BIAS_TRY(hipGetLastError());
    BIAS_TRY(hipMemcpy(out, dout.p, (size_t)m * 8, hipMemcpyDeviceToHost));
    return XMC_OK;
}

int xmc_weight_table_compute(double g, double eps, int base_precision, int table_precision, double* elements,
                             int device) {
    if (!elements) return xmc::set_error(XMC_ERR_ARG, "elements is NULL");
    if (table_precision < 1 || base_precision < 1) return xmc::set_error(XMC_ERR_ARG, "precisions must be >= 1");
    // GetWeightIntegralPrecision, weight_integral.cpp:92-103
    long long n = base_precision;
    if (g > 0.98) n *= 16;
    else if (g > 0.97) n *= 8;
    else if (g > 0.95) n *= 4;
    else if (g > 0.90) n *= 2;
    if (n > 65536) return xmc::set_error(XMC_ERR_LIMIT, "integral precision beyond 65536");
    // entries at x = a + j * step, step = (b - a) / tablePrecision (weight_integral.cpp:50-59)
    const double a = -1.0, b = 1.0, step = (b - a) / table_precision;
    std::vector<double> x((size_t)table_precision + 1);
    for (int j = 0; j <= table_precision; j++) x[j] = a + j * step;
    return xmc_weight_integrals(x.data(), table_precision + 1, g, eps, (int)n, elements, device);
}

}  // extern "C"
